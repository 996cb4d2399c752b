// Mixed-precision contraction: FP32 operands split into three BF16 parts on
// MFMA (gfx950 v_mfma_f32_16x16x32_bf16) with FP32 accumulation; epilogue
// reductions in FP32 within an N-tile and FP64 across N-tiles.
//
// Each FP32 value is written x = h + m + l with h = bf16(x),
// m = bf16(x - h), l = bf16(x - h - m) (24 significant bits, as FP32).  A
// block product keeps the six terms of split order <= 2:
//   hh + hm + mh + hl + mm + lh
// so the product carries FP32 accuracy; the dropped terms are below 2^-26
// relative.  The six MFMAs of a block are issued smallest first so the
// large hh term is added last.
//
// Same tiling and K remapping as the FP64 kernel (contract_f64.hip): each
// lane loads 32 contiguous bytes (8 FP32) of its A row and of each B column
// per 32-wide K-step; lane group g holds k = 8g..8g+7 of both operands.
#include "device_common.h"

namespace gimic {

typedef float v8f __attribute__((ext_vector_type(8)));

struct Split3 {
    v8bf h, m, l;
};

__device__ __forceinline__ Split3 split3(const v8f &x) {
    Split3 s;
    s.h = __builtin_convertvector(x, v8bf);
    const v8f r1 = x - __builtin_convertvector(s.h, v8f);
    s.m = __builtin_convertvector(r1, v8bf);
    const v8f r2 = r1 - __builtin_convertvector(s.m, v8f);
    s.l = __builtin_convertvector(r2, v8bf);
    return s;
}

__device__ __forceinline__ v4f mfma6(const Split3 &a, const Split3 &b, v4f acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.l, b.h, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.m, b.m, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.l, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.m, b.h, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.m, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.h, acc, 0, 0, 0);
    return acc;
}

__global__ __launch_bounds__(256) void contract_bf16x6_kernel(ContractArgs a) {
    __shared__ double qs[kWgRows * kNQ];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int set = blockIdx.z, split = blockIdx.y;
    const size_t row0 = (size_t)blockIdx.x * kWgRows;
    const int lda = a.lda;
    const size_t rows = (size_t)a.rows;
    const size_t matsz = (size_t)a.ncol_pad * lda;
    const size_t gplane = (size_t)a.gplane;
    for (int i = tid; i < kWgRows * kNQ; i += 256) qs[i] = 0.0;
    __syncthreads();

    const float *A0 = a.a32 + (row0 + wave * 32 + c) * lda + 8 * g;
    const float *A1 = A0 + 16 * (size_t)lda;
    const float *Bset = a.b32 + (size_t)set * kNMat * matsz;
    const int nt0 = split * a.tiles_per_split;
    const int nt1 = min(nt0 + a.tiles_per_split, a.ntiles);

    for (int nt = nt0; nt < nt1; nt++) {
        const int col = nt * kNT + c;
        const float *Bc = Bset + (size_t)col * lda + 8 * g;
        v4f acc[2][kNMat];
#pragma unroll
        for (int m = 0; m < kNMat; m++) acc[0][m] = acc[1][m] = v4f{0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < lda; k0 += 32) {
            const Split3 a0 = split3(*(const v8f *)(A0 + k0));
            const Split3 a1 = split3(*(const v8f *)(A1 + k0));
#pragma unroll
            for (int m = 0; m < kNMat; m++) {
                const Split3 b = split3(*(const v8f *)(Bc + m * matsz + k0));
                acc[0][m] = mfma6(a0, b, acc[0][m]);
                acc[1][m] = mfma6(a1, b, acc[1][m]);
            }
        }

        // lane (g, c) holds Y_mat[row 4g + r][col c] (16x16x32 bf16 layout,
        // checked by the MFMA layout self-test)
        const float R[3] = {a.rf32[col], a.rf32[a.ncol_pad + col], a.rf32[2 * a.ncol_pad + col]};
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int rl = wave * 32 + rb * 16 + 4 * g + r;
                const size_t p = row0 + rl;
                const bool valid = p < (size_t)a.npts;
                const float rp[3] = {valid ? (float)a.pts[3 * p] : 0.f,
                                     valid ? (float)a.pts[3 * p + 1] : 0.f,
                                     valid ? (float)a.pts[3 * p + 2] : 0.f};
                const float ph = a.a32[p * lda + col];
                const float G[3] = {a.g32[p * lda + col], a.g32[gplane + p * lda + col],
                                    a.g32[2 * gplane + p * lda + col]};
                const float y0 = acc[rb][0][r];
                float q[kNQ];
                q[0] = y0 * ph;
#pragma unroll
                for (int d = 0; d < 3; d++) q[1 + d] = q[0] * (rp[d] - R[d]);
#pragma unroll
                for (int bb = 0; bb < 3; bb++)
#pragma unroll
                    for (int m = 0; m < 3; m++) q[4 + m + 3 * bb] = acc[rb][1 + bb][r] * G[m];
#pragma unroll
                for (int d = 0; d < 3; d++)
#pragma unroll
                    for (int m = 0; m < 3; m++) q[13 + m + 3 * d] = G[m] * acc[rb][4 + d][r];
#pragma unroll
                for (int off = 1; off < 16; off <<= 1)
#pragma unroll
                    for (int i = 0; i < kNQ; i++) q[i] += __shfl_xor(q[i], off);
#pragma unroll
                for (int i = 0; i < kNQ; i++)
                    if ((i & 15) == c) qs[rl * kNQ + i] += (double)q[i];
            }
        }
    }
    __syncthreads();
    double *out = a.part + ((size_t)set * a.nsplit + split) * kNQ * rows + row0;
    for (int i = tid; i < kWgRows * kNQ; i += 256) {
        const int q = i / kWgRows, rl = i % kWgRows;
        out[(size_t)q * rows + rl] = qs[rl * kNQ + q];
    }
}

hipError_t launch_contract_bf16x6(const ContractArgs &a, hipStream_t s) {
    dim3 grid(a.rows / kWgRows, a.nsplit, a.nmatset);
    hipLaunchKernelGGL(contract_bf16x6_kernel, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace gimic
