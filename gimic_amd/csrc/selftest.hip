// Self-test of the MFMA lane layouts the contraction kernels rely on.
//
// One wavefront runs each builtin twice with A (or B) set to lane+1 and the
// other operand to 1, so every output slot reports which A row / B column
// feeds it.  Outputs (per builtin, 64 lanes x 4 slots each):
//   rows[l*4 + r] = sum over k of (A lane value) for output slot (l, r)
//   cols[l*4 + r] = sum over k of (B lane value) for output slot (l, r)
#include "device_common.h"

namespace gimic {

__global__ void mfma_layout_kernel(double *f64, float *bf16) {
    const int l = threadIdx.x;
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64((double)(l + 1), 1.0, acc, 0, 0, 0);
    for (int r = 0; r < 4; r++) f64[l * 4 + r] = acc[r];
    acc = v4d{0.0, 0.0, 0.0, 0.0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(1.0, (double)(l + 1), acc, 0, 0, 0);
    for (int r = 0; r < 4; r++) f64[256 + l * 4 + r] = acc[r];

    u16x8 lane, one;
    const uint16_t hl = f32_to_bf16_rn((float)(l + 1)), h1 = f32_to_bf16_rn(1.0f);
    for (int e = 0; e < 8; e++) {
        lane[e] = hl;
        one[e] = h1;
    }
    v4f f = {0.f, 0.f, 0.f, 0.f};
    f = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, lane),
                                                 __builtin_bit_cast(v8bf, one), f, 0, 0, 0);
    for (int r = 0; r < 4; r++) bf16[l * 4 + r] = f[r];
    f = v4f{0.f, 0.f, 0.f, 0.f};
    f = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, one),
                                                 __builtin_bit_cast(v8bf, lane), f, 0, 0, 0);
    for (int r = 0; r < 4; r++) bf16[256 + l * 4 + r] = f[r];
}

// 32x32x16 BF16: out[l*16 + e] with A = lane+1, B = 1, then out[1024 + ...]
// with A = 1, B = lane+1
__global__ void mfma32_layout_kernel(float *out) {
    const int l = threadIdx.x;
    typedef float v16f __attribute__((ext_vector_type(16)));
    u16x8 lane, one;
    const uint16_t hl = f32_to_bf16_rn((float)(l + 1)), h1 = f32_to_bf16_rn(1.0f);
    for (int e = 0; e < 8; e++) {
        lane[e] = hl;
        one[e] = h1;
    }
    v16f f;
    for (int e = 0; e < 16; e++) f[e] = 0.f;
    f = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, lane),
                                                 __builtin_bit_cast(v8bf, one), f, 0, 0, 0);
    for (int e = 0; e < 16; e++) out[l * 16 + e] = f[e];
    for (int e = 0; e < 16; e++) f[e] = 0.f;
    f = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, one),
                                                 __builtin_bit_cast(v8bf, lane), f, 0, 0, 0);
    for (int e = 0; e < 16; e++) out[1024 + l * 16 + e] = f[e];
}

hipError_t run_mfma32_layout_probe(float *h_out) {
    float *d = nullptr;
    hipError_t e = hipMalloc(&d, 2048 * sizeof(float));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mfma32_layout_kernel, dim3(1), dim3(64), 0, 0, d);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(h_out, d, 2048 * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return e;
}

hipError_t run_mfma_layout_probe(double *h_f64, float *h_bf16) {
    double *d64 = nullptr;
    float *d32 = nullptr;
    hipError_t e = hipMalloc(&d64, 512 * sizeof(double));
    if (e != hipSuccess) return e;
    e = hipMalloc(&d32, 512 * sizeof(float));
    if (e != hipSuccess) {
        (void)hipFree(d64);
        return e;
    }
    hipLaunchKernelGGL(mfma_layout_kernel, dim3(1), dim3(64), 0, 0, d64, d32);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(h_f64, d64, 512 * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(h_bf16, d32, 512 * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(d64);
    (void)hipFree(d32);
    return e;
}

}  // namespace gimic
