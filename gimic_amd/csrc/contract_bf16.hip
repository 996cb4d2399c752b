// Mixed-precision contraction: split-BF16 operands on MFMA with FP32
// accumulation (gfx950 v_mfma_f32_16x16x32_bf16), epilogue reductions in
// FP32 within an N-tile and FP64 across N-tiles.
//
// Both operands are stored as hi + lo BF16 pairs (x ~ hi + lo to ~16 bits);
// each block product uses the three terms hi*hi + hi*lo + lo*hi (the lo*lo
// term is below the FP32 accumulation error).  Same tiling and K remapping
// as the FP64 kernel (contract_f64.hip); each lane loads 16 contiguous bytes
// (8 BF16) of its A row and of each B column per 32-wide K-step.
#include "device_common.h"

namespace gimic {

__device__ __forceinline__ v4f mfma3(const u16x8 &ah, const u16x8 &al, const u16x8 &bh,
                                     const u16x8 &bl, v4f acc) {
    const v8bf Ah = __builtin_bit_cast(v8bf, ah), Al = __builtin_bit_cast(v8bf, al);
    const v8bf Bh = __builtin_bit_cast(v8bf, bh), Bl = __builtin_bit_cast(v8bf, bl);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al, Bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bh, acc, 0, 0, 0);
    return acc;
}

__global__ __launch_bounds__(256) void contract_bf16x3_kernel(ContractArgs a) {
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

    const size_t aoff0 = (row0 + wave * 32 + c) * lda + 8 * g;
    const size_t aoff1 = aoff0 + 16 * (size_t)lda;
    const size_t bset = (size_t)set * kNMat * matsz;
    const int nt0 = split * a.tiles_per_split;
    const int nt1 = min(nt0 + a.tiles_per_split, a.ntiles);

    for (int nt = nt0; nt < nt1; nt++) {
        const int col = nt * kNT + c;
        const size_t boff = bset + (size_t)col * lda + 8 * g;
        v4f acc[2][kNMat];
#pragma unroll
        for (int m = 0; m < kNMat; m++) acc[0][m] = acc[1][m] = v4f{0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < lda; k0 += 32) {
            const u16x8 ah0 = *(const u16x8 *)(a.ahi + aoff0 + k0);
            const u16x8 al0 = *(const u16x8 *)(a.alo + aoff0 + k0);
            const u16x8 ah1 = *(const u16x8 *)(a.ahi + aoff1 + k0);
            const u16x8 al1 = *(const u16x8 *)(a.alo + aoff1 + k0);
#pragma unroll
            for (int m = 0; m < kNMat; m++) {
                const u16x8 bh = *(const u16x8 *)(a.bhi + boff + m * matsz + k0);
                const u16x8 bl = *(const u16x8 *)(a.blo + boff + m * matsz + k0);
                acc[0][m] = mfma3(ah0, al0, bh, bl, acc[0][m]);
                acc[1][m] = mfma3(ah1, al1, bh, bl, acc[1][m]);
            }
        }

        const float R[3] = {a.rf32[col], a.rf32[a.ncol_pad + col], a.rf32[2 * a.ncol_pad + col]};
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int rl = wave * 32 + rb * 16 + 4 * g + r;
                const size_t p = row0 + rl;
                const float ph = a.phi32[p * lda + col];
                const float G[3] = {a.g32[p * lda + col], a.g32[gplane + p * lda + col],
                                    a.g32[2 * gplane + p * lda + col]};
                const float y0 = acc[rb][0][r];
                float q[kNQ];
                q[0] = y0 * ph;
                q[1] = q[0] * R[0];
                q[2] = q[0] * R[1];
                q[3] = q[0] * R[2];
#pragma unroll
                for (int bb = 0; bb < 3; bb++)
#pragma unroll
                    for (int m = 0; m < 3; m++) q[4 + m + 3 * bb] = acc[rb][1 + bb][r] * G[m];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const float t = R[d] * y0 - acc[rb][4 + d][r];
#pragma unroll
                    for (int m = 0; m < 3; m++) q[13 + m + 3 * d] = G[m] * t;
                }
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

hipError_t launch_contract_bf16x3(const ContractArgs &a, hipStream_t s) {
    dim3 grid(a.rows / kWgRows, a.nsplit, a.nmatset);
    hipLaunchKernelGGL(contract_bf16x3_kernel, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace gimic
