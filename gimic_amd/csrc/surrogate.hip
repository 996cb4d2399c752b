// Learned coordinate surrogate of the J-tensor field: kernels (gfx950).
//
// Network (per molecule): u = (r - centre) / scale;
//   f = [sin(2 pi B u), cos(2 pi B u), u, 0...]   (2F + 3 features, padded to K0)
//   h1..hD = SiLU(W_l h + b_l)                      (width H)
//   y = W_out hD + b_out                            (16 rows: 9 used + 1 log scale)
//   J_i = out_scale_i * y_i * exp(y_9)              (i < 9)
// The learned log amplitude y_9 lets the network follow a field whose size
// changes by orders of magnitude while y_0..8 stay of order one.
// Training: BF16 operands on MFMA (v_mfma_f32_16x16x32_bf16) with FP32
// accumulation, FP32 master weights and Adam.  With t = J / out_scale,
// n = sqrt(|t|^2 + eps^2) and u = t / n, the loss is the mean over points of
//   sum_i (y_i - u_i)^2 + (y_9 - log n)^2
// (direction and log magnitude fitted separately, so every point weighs the
// same whatever the size of its current; the prediction y e^{y_9} then has
// relative error ~ |y - u| + |y_9 - log n|).
// Inference: one fused kernel per 64-point tile, activations in LDS.
//
// The generic GEMM below takes arbitrary element strides for both operands
// (element (m,k) of A at A[m*sam + k*sak], (k,n) of B at B[k*sbk + n*sbn]),
// so the forward product, the input gradient and the weight gradient all use
// it without explicit transposes; tiles are gathered into LDS, zero-padded.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.h"

namespace gimic {

namespace {

constexpr int kGT = 64;   // GEMM tile (M and N)
constexpr int kGK = 32;   // K per step

__device__ __forceinline__ float silu(float z) { return z / (1.f + __expf(-z)); }
// exp of the log amplitude, bounded so that no product overflows FP32
__device__ __forceinline__ float amp(float z) { return __expf(fminf(z, 60.f)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.f / (1.f + __expf(-z));
    return s * (1.f + z * (1.f - s));
}

}  // namespace

// C[M x N] (+)= A[M x K] B[K x N]; bf16 operands, fp32 result.
// grid: (ceil(N/64), ceil(M/64), ksplit); with ksplit > 1 the K range is
// divided and the results are added atomically (C must be zeroed first).
__global__ __launch_bounds__(256) void mlp_gemm_kernel(const __bf16 *A, long long sam, long long sak,
                                                       const __bf16 *B, long long sbk, long long sbn,
                                                       float *C, long long ldc, int M, int N, int K,
                                                       int atomic) {
    __shared__ __align__(16) __bf16 As[kGT][kGK + 8];
    __shared__ __align__(16) __bf16 Bs[kGT][kGK + 8];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int m0 = blockIdx.y * kGT, n0 = blockIdx.x * kGT;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int kchunk = (K + gridDim.z - 1) / gridDim.z;
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    v4f acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += kGK) {
        // gather the 64 x 32 tiles (8 elements per thread each)
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int e = tid + 256 * i;
            const int r = e >> 5, kk = e & 31;
            const int gm = m0 + r, gk = k0 + kk, gn = n0 + r;
            As[r][kk] = (gm < M && gk < ke) ? A[gm * sam + gk * sak] : (__bf16)0.f;
            Bs[r][kk] = (gn < N && gk < ke) ? B[gk * sbk + gn * sbn] : (__bf16)0.f;
        }
        __syncthreads();
        v8bf a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            a[i] = *(const v8bf *)&As[wm + 16 * i + c][8 * g];
            b[i] = *(const v8bf *)&Bs[wn + 16 * i + c][8 * g];
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // lane (g, c) holds rows 4g + r, column c of every 16 x 16 block
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gm = m0 + wm + 16 * i + 4 * g + r, gn = n0 + wn + 16 * j + c;
                if (gm < M && gn < N) {
                    float *dst = C + (long long)gm * ldc + gn;
                    if (atomic) atomicAdd(dst, acc[i][j][r]);
                    else *dst = acc[i][j][r];
                }
            }
}

// Fourier features of npts points into X[p][K0] (bf16) and, optionally, fp32
__global__ void mlp_features_kernel(const double *r, long long npts, const float *Bfreq, int F,
                                    const float *ctr_scale, __bf16 *X, int K0) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    const float inv = 1.f / ctr_scale[3];
    const float u[3] = {(float)(r[3 * p] - ctr_scale[0]) * inv, (float)(r[3 * p + 1] - ctr_scale[1]) * inv,
                        (float)(r[3 * p + 2] - ctr_scale[2]) * inv};
    __bf16 *x = X + p * K0;
    for (int f = 0; f < F; f++) {
        const float t = 6.283185307179586f *
                        (Bfreq[3 * f] * u[0] + Bfreq[3 * f + 1] * u[1] + Bfreq[3 * f + 2] * u[2]);
        float s, co;
        sincosf(t, &s, &co);
        x[f] = (__bf16)s;
        x[F + f] = (__bf16)co;
    }
    x[2 * F] = (__bf16)u[0];
    x[2 * F + 1] = (__bf16)u[1];
    x[2 * F + 2] = (__bf16)u[2];
    for (int k = 2 * F + 3; k < K0; k++) x[k] = (__bf16)0.f;
}

// Z[p][n] += b[n]; H = SiLU(Z) (bf16); Z kept for the backward pass
__global__ void mlp_bias_silu_kernel(float *Z, const float *b, __bf16 *H, long long total, int N) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const float z = Z[i] + b[i % N];
    Z[i] = z;
    H[i] = (__bf16)silu(z);
}

// output layer: Y = Z + b; loss terms and dY (fp32 and bf16):
// dL/dy_i = 2 (y_i - u_i) / npts, dL/dy_9 = 2 (y_9 - log n) / npts
__global__ void mlp_loss_kernel(const float *Z, const float *b, const double *target,
                                const float *out_scale, float eps2, long long npts, int NO,
                                float *dY, __bf16 *dYb, float *loss_acc) {
    __shared__ float red[256];
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (p < npts) {
        float t[9], y[9], tn = eps2;
#pragma unroll
        for (int i = 0; i < 9; i++) {
            t[i] = (float)target[9 * p + i] / out_scale[i];
            y[i] = Z[p * NO + i] + b[i];
            tn += t[i] * t[i];
        }
        const float inv = rsqrtf(tn);
        const float dz = (Z[p * NO + 9] + b[9]) - 0.5f * __logf(tn);
        const float s = 2.f / (float)npts;
        l = dz * dz;
#pragma unroll
        for (int i = 0; i < 9; i++) {
            const float d = y[i] - t[i] * inv;
            l += d * d;
            dY[p * NO + i] = s * d;
            dYb[p * NO + i] = (__bf16)(s * d);
        }
        dY[p * NO + 9] = s * dz;
        dYb[p * NO + 9] = (__bf16)(s * dz);
        for (int i = 10; i < NO; i++) {
            dY[p * NO + i] = 0.f;
            dYb[p * NO + i] = (__bf16)0.f;
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss_acc, red[0] / (float)npts);
}

// dZ = dH * SiLU'(Z): fp32 copy (bias gradient) and bf16 copy (GEMMs)
__global__ void mlp_dsilu_kernel(const float *dH, const float *Z, float *dZ, __bf16 *dZb,
                                 long long total) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const float d = dH[i] * dsilu(Z[i]);
    dZ[i] = d;
    dZb[i] = (__bf16)d;
}

// db[n] += sum_p dZ[p][n] (column sums, atomic across blocks)
__global__ void mlp_colsum_kernel(const float *dZ, long long npts, int N, float *db) {
    const int n = blockIdx.x;
    float s = 0.f;
    for (long long p = blockIdx.y * blockDim.x + threadIdx.x; p < npts;
         p += (long long)gridDim.y * blockDim.x)
        s += dZ[p * N + n];
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(db + n, red[0]);
}

// Adam on the fp32 master parameters; refresh the bf16 weight copies
__global__ void mlp_adam_kernel(float *p, const float *grad, float *m, float *v, long long n,
                                float lr, float b1, float b2, float eps, float bc1, float bc2,
                                __bf16 *pb) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gr = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * gr;
    const float vi = b2 * v[i] + (1.f - b2) * gr * gr;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    p[i] = pi;
    pb[i] = (__bf16)pi;
}

__global__ void mlp_cast_kernel(const float *p, __bf16 *pb, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pb[i] = (__bf16)p[i];
}

// final output: J[p][i] = (Z[p][i] + b[i]) exp(Z[p][9] + b[9]) out_scale[i] (fp64)
__global__ void mlp_output_kernel(const float *Z, const float *b, const float *out_scale,
                                  long long npts, int NO, double *J) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    const float e = amp(Z[p * NO + 9] + b[9]);
#pragma unroll
    for (int i = 0; i < 9; i++) J[9 * p + i] = (double)((Z[p * NO + i] + b[i]) * e * out_scale[i]);
}

// ---- fused inference: one 64-point tile per workgroup ----
// weights bf16, layout W_l[n][k]; activations in LDS [64][width + 8]
template <int K0, int H, int D>
__global__ __launch_bounds__(256) void mlp_infer_kernel(const double *r, long long npts,
                                                        const float *Bfreq, int F,
                                                        const float *ctr_scale,
                                                        const __bf16 *const *W, const float *const *bias,
                                                        const float *out_scale, double *J) {
    constexpr int P = 64;
    constexpr int LD = (K0 > H ? K0 : H) + 8;
    __shared__ __align__(16) __bf16 act[2][P][LD];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const long long p0 = (long long)blockIdx.x * P;
    // features: 4 threads per point
    {
        const int pl = tid >> 2, part = tid & 3;
        const long long p = p0 + pl;
        const float inv = 1.f / ctr_scale[3];
        float u[3] = {0.f, 0.f, 0.f};
        if (p < npts) {
            u[0] = (float)(r[3 * p] - ctr_scale[0]) * inv;
            u[1] = (float)(r[3 * p + 1] - ctr_scale[1]) * inv;
            u[2] = (float)(r[3 * p + 2] - ctr_scale[2]) * inv;
        }
        for (int f = part; f < F; f += 4) {
            const float t = 6.283185307179586f *
                            (Bfreq[3 * f] * u[0] + Bfreq[3 * f + 1] * u[1] + Bfreq[3 * f + 2] * u[2]);
            float s, co;
            sincosf(t, &s, &co);
            act[0][pl][f] = (__bf16)s;
            act[0][pl][F + f] = (__bf16)co;
        }
        if (part == 0) {
            act[0][pl][2 * F] = (__bf16)u[0];
            act[0][pl][2 * F + 1] = (__bf16)u[1];
            act[0][pl][2 * F + 2] = (__bf16)u[2];
            for (int k = 2 * F + 3; k < K0; k++) act[0][pl][k] = (__bf16)0.f;
        }
    }
    __syncthreads();
    int cur = 0;
    // hidden layers: each wave computes 16 rows x H columns
#pragma unroll 1
    for (int l = 0; l < D; l++) {
        const int K = l == 0 ? K0 : H;
        const __bf16 *Wl = W[l];
        const float *bl = bias[l];
        v4f acc[H / 16];
#pragma unroll
        for (int j = 0; j < H / 16; j++) acc[j] = v4f{0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < K; k0 += 32) {
            const v8bf a = *(const v8bf *)&act[cur][wave * 16 + c][k0 + 8 * g];
#pragma unroll
            for (int j = 0; j < H / 16; j++) {
                const v8bf b = *(const v8bf *)(Wl + (long long)(16 * j + c) * K + k0 + 8 * g);
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
            }
        }
#pragma unroll
        for (int j = 0; j < H / 16; j++)
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const int n = 16 * j + c;
                act[cur ^ 1][wave * 16 + 4 * g + rr][n] = (__bf16)silu(acc[j][rr] + bl[n]);
            }
        __syncthreads();
        cur ^= 1;
    }
    // output layer (16 columns: 9 used + the log amplitude in column 9)
    {
        const __bf16 *Wo = W[D];
        const float *bo = bias[D];
        v4f acc = {0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < H; k0 += 32) {
            const v8bf a = *(const v8bf *)&act[cur][wave * 16 + c][k0 + 8 * g];
            const v8bf b = *(const v8bf *)(Wo + (long long)c * H + k0 + 8 * g);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            const long long p = p0 + wave * 16 + 4 * g + rr;
            const float y = acc[rr] + bo[c];
            // lane (g, 9) holds the log amplitude of the same point
            const float e = amp(__shfl(y, (lane & ~15) | 9));
            if (c < 9 && p < npts) J[9 * p + c] = (double)(y * e * out_scale[c]);
        }
    }
}

// ---------------------------------------------------------------- launchers

hipError_t mlp_gemm(const __bf16 *A, long long sam, long long sak, const __bf16 *B, long long sbk,
                    long long sbn, float *C, long long ldc, int M, int N, int K, int ksplit,
                    hipStream_t s) {
    dim3 grid((N + kGT - 1) / kGT, (M + kGT - 1) / kGT, ksplit);
    hipLaunchKernelGGL(mlp_gemm_kernel, grid, dim3(256), 0, s, A, sam, sak, B, sbk, sbn, C, ldc,
                       M, N, K, ksplit > 1 ? 1 : 0);
    return hipGetLastError();
}

static inline unsigned blocks(long long n, int t) { return (unsigned)((n + t - 1) / t); }

hipError_t mlp_features(const double *r, long long npts, const float *Bfreq, int F,
                        const float *cs, __bf16 *X, int K0, hipStream_t s) {
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_features_kernel, dim3(blocks(npts, 256)), dim3(256), 0, s, r, npts,
                       Bfreq, F, cs, X, K0);
    return hipGetLastError();
}

hipError_t mlp_bias_silu(float *Z, const float *b, __bf16 *H, long long total, int N,
                         hipStream_t s) {
    hipLaunchKernelGGL(mlp_bias_silu_kernel, dim3(blocks(total, 256)), dim3(256), 0, s, Z, b, H,
                       total, N);
    return hipGetLastError();
}

hipError_t mlp_loss(const float *Z, const float *b, const double *target, const float *out_scale,
                    float eps2, long long npts, int NO, float *dY, __bf16 *dYb, float *loss,
                    hipStream_t s) {
    hipLaunchKernelGGL(mlp_loss_kernel, dim3(blocks(npts, 256)), dim3(256), 0, s, Z, b, target,
                       out_scale, eps2, npts, NO, dY, dYb, loss);
    return hipGetLastError();
}

hipError_t mlp_dsilu(const float *dH, const float *Z, float *dZ, __bf16 *dZb, long long total,
                     hipStream_t s) {
    hipLaunchKernelGGL(mlp_dsilu_kernel, dim3(blocks(total, 256)), dim3(256), 0, s, dH, Z, dZ, dZb,
                       total);
    return hipGetLastError();
}

hipError_t mlp_colsum(const float *dZ, long long npts, int N, float *db, hipStream_t s) {
    hipLaunchKernelGGL(mlp_colsum_kernel, dim3(N, 32), dim3(256), 0, s, dZ, npts, N, db);
    return hipGetLastError();
}

hipError_t mlp_adam(float *p, const float *grad, float *m, float *v, long long n, float lr,
                    float b1, float b2, float eps, float bc1, float bc2, __bf16 *pb,
                    hipStream_t s) {
    hipLaunchKernelGGL(mlp_adam_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, p, grad, m, v, n,
                       lr, b1, b2, eps, bc1, bc2, pb);
    return hipGetLastError();
}

hipError_t mlp_cast(const float *p, __bf16 *pb, long long n, hipStream_t s) {
    hipLaunchKernelGGL(mlp_cast_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, p, pb, n);
    return hipGetLastError();
}

hipError_t mlp_output(const float *Z, const float *b, const float *out_scale, long long npts,
                      int NO, double *J, hipStream_t s) {
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_output_kernel, dim3(blocks(npts, 256)), dim3(256), 0, s, Z, b,
                       out_scale, npts, NO, J);
    return hipGetLastError();
}

hipError_t mlp_infer(const double *r, long long npts, const float *Bfreq, int F, const float *cs,
                     const __bf16 *const *W, const float *const *bias, const float *out_scale,
                     double *J, hipStream_t s) {
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL((mlp_infer_kernel<160, 128, 4>), dim3(blocks(npts, 64)), dim3(256), 0, s, r,
                       npts, Bfreq, F, cs, W, bias, out_scale, J);
    return hipGetLastError();
}

}  // namespace gimic
