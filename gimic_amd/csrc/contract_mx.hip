// Mixed-precision contraction, LDS-tiled (gfx950).
//
// Same arithmetic as contract_bf16x6_kernel (FP32 operands split into three
// BF16 parts, six v_mfma_f32_16x16x32_bf16 products per block, FP32
// accumulation, FP64 partial sums across N-tiles) with a larger tile and the
// right operand staged through LDS:
//   * a workgroup (4 wavefronts) owns 128 point rows; each wavefront keeps
//     32 rows x 224 columns (7 matrices x 32 functions) of accumulators,
//     2 x 14 MFMA blocks;
//   * per 32-wide K-step the 224 x 32 FP32 slice of the right operand is
//     read once per workgroup (coalesced 16-byte loads), split into BF16
//     planes and written to one of two LDS buffers while the matrix cores
//     work on the other;
//   * each wavefront loads its own A fragments (8 FP32 per lane and row
//     block) one K-step ahead and splits them in registers.
// LDS layout of a B buffer: [plane][k-group g][column][8 k] (16 bytes per
// column and k-group) so the 16 lanes of a k-group read 256 contiguous bytes.
// B fragments are read one column block ahead and scheduling barriers keep
// the compiler from hoisting every read (and the epilogue loads) at once,
// which bounds register use to one wavefront per SIMD without spills.
#include "device_common.h"

namespace gimic {

namespace {

typedef float v8f __attribute__((ext_vector_type(8)));

constexpr int kRows = 128;            // point rows per workgroup
constexpr int kNu = 32;               // functions per matrix per N-tile
constexpr int kCols = kNMat * kNu;    // 224 columns per N-tile
constexpr int kCB = kCols / 16;       // 14 column blocks
constexpr int kRB = 2;                // row blocks per wavefront
constexpr int kBufElems = 3 * 4 * kCols * 8;             // bf16 per buffer
constexpr size_t kLdsB = 2 * (size_t)kBufElems * 2;       // bytes, two buffers
constexpr size_t kLdsQ = (size_t)kRows * kNQ * 8;         // FP64 partial sums
constexpr int kChunks = kCols * 8 / 256;                  // 16-byte B chunks per thread

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

__device__ __forceinline__ v4f mfma6(const Split3 &a, const v8bf &bh, const v8bf &bm,
                                     const v8bf &bl, v4f acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.l, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.m, bm, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.m, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, bm, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, bh, acc, 0, 0, 0);
    return acc;
}

typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

// split 4 FP32 of one column into the three planes of an LDS buffer
__device__ __forceinline__ void stage_chunk(__bf16 *buf, int col, int kq, const v4f &x) {
    const v4bf h = __builtin_convertvector(x, v4bf);
    const v4f r1 = x - __builtin_convertvector(h, v4f);
    const v4bf m = __builtin_convertvector(r1, v4bf);
    const v4f r2 = r1 - __builtin_convertvector(m, v4f);
    const v4bf l = __builtin_convertvector(r2, v4bf);
    const int g = kq >> 1, half = kq & 1;
    const int off = (g * kCols + col) * 8 + half * 4;
    *(v4bf *)(buf + 0 * 4 * kCols * 8 + off) = h;
    *(v4bf *)(buf + 1 * 4 * kCols * 8 + off) = m;
    *(v4bf *)(buf + 2 * 4 * kCols * 8 + off) = l;
}

}  // namespace

__global__ __launch_bounds__(256, 1) void contract_mx_kernel(ContractArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __bf16 *bbuf = (__bf16 *)smem;
    double *qs = (double *)(smem + kLdsB);

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const int set = blockIdx.z, split = blockIdx.y;
    const size_t row0 = (size_t)blockIdx.x * kRows;
    const int lda = a.lda;
    const size_t rows = (size_t)a.rows;
    const size_t matsz = (size_t)a.ncol_pad * lda;
    const size_t gplane = (size_t)a.gplane;
    const int nks = lda / 32;
    for (int i = tid; i < kRows * kNQ; i += 256) qs[i] = 0.0;

    const float *Aw = a.a32 + (row0 + wave * 32 + c) * lda + 8 * g;
    const float *Bset = a.b32 + (size_t)set * kNMat * matsz;
    const int nt0 = split * a.tiles_per_split;
    const int nt1 = min(nt0 + a.tiles_per_split, a.ntiles);

    // this thread's B chunks: column col = id/8 (matrix col/32, function
    // nt*32 + col%32), k-quad kq = id%8 (k = 4kq..4kq+3 of the K-step)
    int ccol[kChunks], ckq[kChunks];
    int coff[kChunks];
#pragma unroll
    for (int i = 0; i < kChunks; i++) {
        const int id = tid + 256 * i;
        ccol[i] = id >> 3;
        ckq[i] = id & 7;
        coff[i] = (ccol[i] / kNu) * (int)matsz + (ccol[i] % kNu) * lda + 4 * ckq[i];
    }

    for (int nt = nt0; nt < nt1; nt++) {
        const float *Bnt = Bset + (size_t)nt * kNu * lda;
        v4f acc[kRB][kCB];
#pragma unroll
        for (int rb = 0; rb < kRB; rb++)
#pragma unroll
            for (int cb = 0; cb < kCB; cb++) acc[rb][cb] = v4f{0.f, 0.f, 0.f, 0.f};

        // stage K-step 0
        {
            v4f x[kChunks];
#pragma unroll
            for (int i = 0; i < kChunks; i++) x[i] = *(const v4f *)(Bnt + coff[i]);
#pragma unroll
            for (int i = 0; i < kChunks; i++) stage_chunk(bbuf, ccol[i], ckq[i], x[i]);
        }
        v8f an[kRB];
#pragma unroll
        for (int rb = 0; rb < kRB; rb++) an[rb] = *(const v8f *)(Aw + (size_t)rb * 16 * lda);
        __syncthreads();

        for (int ks = 0; ks < nks; ks++) {
            const bool more = ks + 1 < nks;
            const int kn = more ? (ks + 1) * 32 : ks * 32;
            Split3 as[kRB];
#pragma unroll
            for (int rb = 0; rb < kRB; rb++) as[rb] = split3(an[rb]);
            // prefetch the next K-step (A fragments and the B slice)
            v4f x[kChunks];
#pragma unroll
            for (int i = 0; i < kChunks; i++) x[i] = *(const v4f *)(Bnt + coff[i] + kn);
#pragma unroll
            for (int rb = 0; rb < kRB; rb++)
                an[rb] = *(const v8f *)(Aw + (size_t)rb * 16 * lda + kn);

            const __bf16 *cur = bbuf + (size_t)(ks & 1) * kBufElems;
            v8bf bh[2], bm[2], bl[2];
            {
                const int off = (g * kCols + c) * 8;
                bh[0] = *(const v8bf *)(cur + 0 * 4 * kCols * 8 + off);
                bm[0] = *(const v8bf *)(cur + 1 * 4 * kCols * 8 + off);
                bl[0] = *(const v8bf *)(cur + 2 * 4 * kCols * 8 + off);
            }
#pragma unroll
            for (int cb = 0; cb < kCB; cb++) {
                const int sl = cb & 1;
                if (cb + 1 < kCB) {
                    const int off = (g * kCols + (cb + 1) * 16 + c) * 8;
                    bh[sl ^ 1] = *(const v8bf *)(cur + 0 * 4 * kCols * 8 + off);
                    bm[sl ^ 1] = *(const v8bf *)(cur + 1 * 4 * kCols * 8 + off);
                    bl[sl ^ 1] = *(const v8bf *)(cur + 2 * 4 * kCols * 8 + off);
                }
#pragma unroll
                for (int rb = 0; rb < kRB; rb++)
                    acc[rb][cb] = mfma6(as[rb], bh[sl], bm[sl], bl[sl], acc[rb][cb]);
                __builtin_amdgcn_sched_barrier(0);
            }
            if (more) {
                __bf16 *nxt = bbuf + (size_t)((ks + 1) & 1) * kBufElems;
#pragma unroll
                for (int i = 0; i < kChunks; i++) stage_chunk(nxt, ccol[i], ckq[i], x[i]);
            }
            __syncthreads();
        }

        // epilogue: lane (g, c) holds Y[row 4g + r][col c] of every block;
        // column block cb = 2 * matrix + half covers functions nt*32 + 16*half + c
        float R[2][3];
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int d = 0; d < 3; d++) R[h][d] = a.rf32[(size_t)d * a.ncol_pad + nt * kNu + 16 * h + c];
#pragma unroll
        for (int rb = 0; rb < kRB; rb++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int rl = wave * 32 + rb * 16 + 4 * g + r;
                const size_t p = row0 + rl;
                const bool valid = p < (size_t)a.npts;
                const float rp[3] = {valid ? (float)a.pts[3 * p] : 0.f,
                                     valid ? (float)a.pts[3 * p + 1] : 0.f,
                                     valid ? (float)a.pts[3 * p + 2] : 0.f};
                float q[kNQ];
#pragma unroll
                for (int i = 0; i < kNQ; i++) q[i] = 0.f;
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const size_t e = p * lda + nt * kNu + 16 * h + c;
                    const float ph = a.a32[e];
                    const float G[3] = {a.g32[e], a.g32[gplane + e], a.g32[2 * gplane + e]};
                    const float y0 = acc[rb][h][r];
                    const float t0 = y0 * ph;
                    q[0] += t0;
#pragma unroll
                    for (int d = 0; d < 3; d++) q[1 + d] += t0 * (rp[d] - R[h][d]);
#pragma unroll
                    for (int bb = 0; bb < 3; bb++)
#pragma unroll
                        for (int m = 0; m < 3; m++) q[4 + m + 3 * bb] += acc[rb][2 * (1 + bb) + h][r] * G[m];
#pragma unroll
                    for (int d = 0; d < 3; d++)
#pragma unroll
                        for (int m = 0; m < 3; m++) q[13 + m + 3 * d] += G[m] * acc[rb][2 * (4 + d) + h][r];
                }
#pragma unroll
                for (int off = 1; off < 16; off <<= 1)
#pragma unroll
                    for (int i = 0; i < kNQ; i++) q[i] += __shfl_xor(q[i], off);
#pragma unroll
                for (int i = 0; i < kNQ; i++)
                    if ((i & 15) == c) qs[rl * kNQ + i] += (double)q[i];
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    __syncthreads();
    double *out = a.part + ((size_t)set * a.nsplit + split) * kNQ * rows + row0;
    for (int i = tid; i < kRows * kNQ; i += 256) {
        const int q = i / kRows, rl = i % kRows;
        out[(size_t)q * rows + rl] = qs[rl * kNQ + q];
    }
}

hipError_t launch_contract_mx(const ContractArgs &a, hipStream_t s) {
    static bool attr = false;
    const size_t lds = kLdsB + kLdsQ;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)contract_mx_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr = true;
    }
    dim3 grid(a.rows / kRows, a.nsplit, a.nmatset);
    hipLaunchKernelGGL(contract_mx_kernel, grid, dim3(256), lds, s, a);
    return hipGetLastError();
}

int contract_mx_rows() { return kRows; }
int contract_mx_cols() { return kNu; }

}  // namespace gimic
