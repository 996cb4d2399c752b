// Mixed-precision contraction on 32x32 MFMA blocks (gfx950).
//
// Same arithmetic and operand formats as contract_mx2.hip (BF16 parts
// h + m + l of FP32 values, the six products of split order <= 2, FP32
// accumulation, FP64 partial sums across N-tiles; A split by the basis
// kernel as [row][k/8][3][8], B as 44 KiB LDS-image blocks streamed with
// direct global-to-LDS loads), but on v_mfma_f32_32x32x16_bf16.  A 32x32x16
// instruction does twice the work of a 16x16x32 one for the same operand
// registers, and with one wavefront per SIMD it sustains ~89% of the BF16
// peak where 16x16x32 stops near 60% (measured on MI355X with an
// independent-accumulator loop).
//
// Tile: 128 point rows per workgroup (4 wavefronts x 32 rows); per N-tile of
// 32 functions a wavefront holds one 32x32 accumulator block per matrix
// (7 x 16 AGPRs).  Lane (c, q) = (lane % 32, lane / 32) holds A row c and B
// column c for k-group q of each 16-wide half of the K-step, and output rows
// 8 (e / 4) + 4 q + e % 4, column c, in accumulator element e (checked by
// the MFMA layout self-test).
//
// K-step order: for each half h and each B part pb = l, m, h (smallest
// first) the seven matrices' B fragments are read from LDS (one ds_read_b128
// per matrix, the next group read ahead), then every A part pa with
// pa + pb <= 2 multiplies all seven: consecutive MFMAs use different
// accumulators, so no instruction waits for its predecessor's result.
//
// Epilogue: the 22 per-point partial sums are sums over the 32 columns held
// by the 32 lanes of a half-wavefront; a transposing butterfly (each step
// exchanges half of the remaining values and adds) leaves sum i in lane i
// after 31 shuffles instead of 22 x 5.
#include "device_common.h"

namespace gimic {

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void lds_void;

constexpr int kRows = 128;
constexpr int kNu = 32;
constexpr int kCols = kMx2Cols;                // 224 = 7 x 32
constexpr int kBlk = kMx2BlockElems;           // BF16 per block / LDS buffer
constexpr int kPlane = 4 * kCols * 8;          // BF16 per part plane of a block
constexpr int kPieces = kBlk * 2 / 1024;       // 1 KiB pieces per block (44)
constexpr int kPiecesPerWave = kPieces / 4;    // 11
constexpr size_t kBufBytes = (size_t)kBlk * 2;
constexpr size_t kLdsQ = (size_t)kRows * kNQ * 8;
constexpr int kWaitAll = 0;
constexpr int kWaitAllButBlock = (kPiecesPerWave & 15) | ((kPiecesPerWave >> 4) << 14) | (7 << 4);

__device__ __forceinline__ void stage_piece(const __bf16 *blk, unsigned char *buf, int wave,
                                            int lane, int i) {
    const int piece = wave + 4 * i;
    const unsigned char *src = (const unsigned char *)blk + piece * 1024 + lane * 16;
    __builtin_amdgcn_global_load_lds((const void *)src, (lds_void *)(buf + piece * 1024), 16, 0,
                                     0);
}

__device__ __forceinline__ void stage_block(const __bf16 *blk, unsigned char *buf, int wave,
                                            int lane) {
#pragma unroll
    for (int i = 0; i < kPiecesPerWave; i++) stage_piece(blk, buf, wave, lane, i);
}

// seven B fragments (one per matrix) of part pb, half h
__device__ __forceinline__ void read_group(v8bf (&bf)[7], const __bf16 *cur, int pb, int h, int q,
                                           int c) {
#pragma unroll
    for (int m = 0; m < 7; m++)
        bf[m] = *(const v8bf *)(cur + (2 * h + q) * kMx2GroupElems + pb * (kCols * 8) +
                                (m * 32 + c) * 8);
}

// one butterfly step: lanes with bit W set keep the upper W values, the
// others the lower W, and each adds the partner's copy of what it keeps
// (W is a template argument so every index is a compile-time constant: a
// runtime step width turns v[] into a dynamically indexed array)
template <int W>
__device__ __forceinline__ void butterfly_step(float (&v)[32], int c) {
    const bool up = (c & W) != 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const float lo = v[i], hi = v[i + W];
        const float send = up ? lo : hi;
        const float keep = up ? hi : lo;
        v[i] = keep + __shfl_xor(send, W);
    }
}

// v[0..31] summed over the 32 lanes of each half-wavefront; returns sum c in lane c
__device__ __forceinline__ float butterfly32(float (&v)[32], int c) {
    butterfly_step<16>(v, c);
    butterfly_step<8>(v, c);
    butterfly_step<4>(v, c);
    butterfly_step<2>(v, c);
    butterfly_step<1>(v, c);
    return v[0];
}

}  // namespace

template <int NBUF>
__global__ __launch_bounds__(256, 1) void contract_mx3_kernel(ContractArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned char *bbuf = smem;
    double *qs = (double *)(smem + NBUF * kBufBytes);

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c = lane & 31, q = lane >> 5;
    // XCD-aware order as in contract_mx2.hip
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int groups = (a.rows / kRows + 7) >> 3;
    const int per_set = groups * a.nsplit;
    const int set = slot / per_set;
    const int split = slot % a.nsplit;
    const int rtile = ((slot % per_set) / a.nsplit) * 8 + xcd;
    if (rtile >= a.rows / kRows) return;
    const size_t row0 = (size_t)rtile * kRows;
    const int lda = a.lda;
    const int nks = lda / 32;
    const size_t rows = (size_t)a.rows;
    const size_t gplane = (size_t)a.gplane;
    const size_t arow = 3 * (size_t)lda;
    for (int i = tid; i < kRows * kNQ; i += 256) qs[i] = 0.0;

    // A row wave*32 + c; k-group q of each half
    const __bf16 *Aw = a.as3 + (row0 + wave * 32 + c) * arow + q * 24;
    const __bf16 *Bset = a.bs3 + (size_t)set * a.ntiles * nks * kMx2StepElems;
    const int nt0 = split * a.tiles_per_split;
    const int nt1 = min(nt0 + a.tiles_per_split, a.ntiles);

    for (int nt = nt0; nt < nt1; nt++) {
        const __bf16 *Bnt = Bset + (size_t)nt * nks * kMx2StepElems;
        v16f acc[7];
#pragma unroll
        for (int m = 0; m < 7; m++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[m][e] = 0.f;

        stage_block(Bnt, bbuf, wave, lane);
        if (NBUF == 3 && nks > 1) stage_block(Bnt + kMx2StepElems, bbuf + kBufBytes, wave, lane);
        v8bf an[2][3];
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int p = 0; p < 3; p++) an[h][p] = *(const v8bf *)(Aw + 2 * h * 24 + p * 8);
        __builtin_amdgcn_s_waitcnt(kWaitAll);
        __syncthreads();

        for (int ks = 0; ks < nks; ks++) {
            v8bf as[2][3];
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int p = 0; p < 3; p++) as[h][p] = an[h][p];
            // vector-memory instructions spread over the MFMA groups (see
            // contract_mx2.hip): the 6 A loads three per group in groups 0
            // and 1, the copy's 11 pieces after them (NBUF = 3: groups 2 to
            // 5) or two per group (NBUF = 2)
            const bool more = ks + 1 < nks;
            const int ahead = ks + NBUF - 1;
            const bool copy = NBUF == 3 || ahead < nks;
            const __bf16 *cblk = Bnt + (size_t)min(ahead, nks - 1) * kMx2StepElems;
            unsigned char *cbuf = bbuf + (size_t)(ahead % NBUF) * kBufBytes;
            const __bf16 *cur = (const __bf16 *)(bbuf + (size_t)(ks % NBUF) * kBufBytes);

            v8bf bf[2][7];
            read_group(bf[0], cur, 2, 0, q, c);
#pragma unroll
            for (int grp = 0; grp < 6; grp++) {
                const int h = grp / 3, pb = 2 - grp % 3;
                if (grp + 1 < 6) read_group(bf[(grp + 1) & 1], cur, 2 - (grp + 1) % 3, (grp + 1) / 3, q, c);
#pragma unroll
                for (int j = 3 * grp; j < 3 * grp + 3; j++)
                    if (more && j < 6)
                        an[j / 3][j % 3] = *(const v8bf *)(Aw + ((size_t)(ks + 1) * 4 + 2 * (j / 3)) * 24 +
                                                          (j % 3) * 8);
                if (copy) {
                    if (NBUF == 3) {
                        // pieces 0-2, 3-5, 6-8, 9-10 in groups 2..5
#pragma unroll
                        for (int i = 3 * (grp - 2); i < 3 * (grp - 2) + 3; i++)
                            if (grp >= 2 && i < kPiecesPerWave) stage_piece(cblk, cbuf, wave, lane, i);
                    } else {
                        if (2 * grp < kPiecesPerWave) stage_piece(cblk, cbuf, wave, lane, 2 * grp);
                        if (2 * grp + 1 < kPiecesPerWave) stage_piece(cblk, cbuf, wave, lane, 2 * grp + 1);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int pa = 2 - pb; pa >= 0; pa--)
#pragma unroll
                    for (int m = 0; m < 7; m++)
                        acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as[h][pa], bf[grp & 1][m],
                                                                         acc[m], 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
            }
            if (NBUF == 3) __builtin_amdgcn_s_waitcnt(kWaitAllButBlock);
            else __builtin_amdgcn_s_waitcnt(kWaitAll);
            __builtin_amdgcn_s_barrier();
        }
        if (NBUF == 3) {
            __builtin_amdgcn_s_waitcnt(kWaitAll);
            __builtin_amdgcn_s_barrier();
        }

        // epilogue: element e of acc[m] is Y_m[row 8(e/4) + 4q + e%4][function nt*32 + c]
        const int col = nt * kNu + c;
        const float R[3] = {a.rf32[col], a.rf32[(size_t)a.ncol_pad + col],
                            a.rf32[2 * (size_t)a.ncol_pad + col]};
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const int rl = wave * 32 + 8 * (e >> 2) + 4 * q + (e & 3);
            const size_t p = row0 + rl;
            const bool valid = p < (size_t)a.npts;
            const float rp[3] = {valid ? (float)a.pts[3 * p] : 0.f,
                                 valid ? (float)a.pts[3 * p + 1] : 0.f,
                                 valid ? (float)a.pts[3 * p + 2] : 0.f};
            const __bf16 *pe = a.as3 + p * arow + (col >> 3) * 24 + (col & 7);
            const float ph = ((float)pe[0] + (float)pe[8]) + (float)pe[16];
            const size_t ei = p * lda + col;
            const float G[3] = {a.g32[ei], a.g32[gplane + ei], a.g32[2 * gplane + ei]};
            float v[32];
            const float t0 = acc[0][e] * ph;
            v[0] = t0;
#pragma unroll
            for (int d = 0; d < 3; d++) v[1 + d] = t0 * (rp[d] - R[d]);
#pragma unroll
            for (int bb = 0; bb < 3; bb++)
#pragma unroll
                for (int m = 0; m < 3; m++) v[4 + m + 3 * bb] = acc[1 + bb][e] * G[m];
#pragma unroll
            for (int d = 0; d < 3; d++)
#pragma unroll
                for (int m = 0; m < 3; m++) v[13 + m + 3 * d] = G[m] * acc[4 + d][e];
#pragma unroll
            for (int i = kNQ; i < 32; i++) v[i] = 0.f;
            const float s = butterfly32(v, c);
            if (c < kNQ) qs[rl * kNQ + c] += (double)s;
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    __syncthreads();
    double *out = a.part + ((size_t)set * a.nsplit + split) * kNQ * rows + row0;
    for (int i = tid; i < kRows * kNQ; i += 256) {
        const int qq = i / kRows, rl = i % kRows;
        out[(size_t)qq * rows + rl] = qs[rl * kNQ + qq];
    }
}

// Pipelined form (three LDS buffers).  B fragments are read two MFMA groups
// ahead from a ring of three register sets; the K-step's synchronisation
// sits after its second group (wait until this wavefront's copy of block
// ks+1 has landed, then a bare barrier), so the last two groups of K-step ks
// already read the first two groups of K-step ks+1 and no LDS latency is
// exposed at K-step boundaries.  The copy of block ks+2 is issued after
// that barrier, into the buffer that held block ks-1, which every wavefront
// has finished reading by then.  All A loads and copies are issued on every
// K-step (indices clamped at the end) so the count of outstanding loads is
// the same on every iteration.
__device__ __forceinline__ int grp_pb(int grp) { return 2 - grp % 3; }
__device__ __forceinline__ int grp_h(int grp) { return grp / 3; }

__global__ __launch_bounds__(256, 1) void contract_mx3p_kernel(ContractArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned char *bbuf = smem;
    double *qs = (double *)(smem + 3 * kBufBytes);

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c = lane & 31, q = lane >> 5;
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int groups = (a.rows / kRows + 7) >> 3;
    const int per_set = groups * a.nsplit;
    const int set = slot / per_set;
    const int split = slot % a.nsplit;
    const int rtile = ((slot % per_set) / a.nsplit) * 8 + xcd;
    if (rtile >= a.rows / kRows) return;
    const size_t row0 = (size_t)rtile * kRows;
    const int lda = a.lda;
    const int nks = lda / 32;
    const size_t rows = (size_t)a.rows;
    const size_t gplane = (size_t)a.gplane;
    const size_t arow = 3 * (size_t)lda;
    for (int i = tid; i < kRows * kNQ; i += 256) qs[i] = 0.0;

    const __bf16 *Aw = a.as3 + (row0 + wave * 32 + c) * arow + q * 24;
    const __bf16 *Bset = a.bs3 + (size_t)set * a.ntiles * nks * kMx2StepElems;
    const int nt0 = split * a.tiles_per_split;
    const int nt1 = min(nt0 + a.tiles_per_split, a.ntiles);
    // vmcnt(6): all but the six A loads issued after the last copy
    constexpr int kWaitAllButA = 6 | (7 << 4) | (15 << 8);

    for (int nt = nt0; nt < nt1; nt++) {
        const __bf16 *Bnt = Bset + (size_t)nt * nks * kMx2StepElems;
        v16f acc[7];
#pragma unroll
        for (int m = 0; m < 7; m++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[m][e] = 0.f;

        stage_block(Bnt, bbuf, wave, lane);
        stage_block(Bnt + (size_t)min(1, nks - 1) * kMx2StepElems, bbuf + kBufBytes, wave, lane);
        v8bf an[2][3];
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int p = 0; p < 3; p++) an[h][p] = *(const v8bf *)(Aw + 2 * h * 24 + p * 8);
        __builtin_amdgcn_s_waitcnt(kWaitAll);
        __syncthreads();
        v8bf bf[3][7];
        read_group(bf[0], (const __bf16 *)bbuf, grp_pb(0), grp_h(0), q, c);
        read_group(bf[1], (const __bf16 *)bbuf, grp_pb(1), grp_h(1), q, c);

        for (int ks = 0; ks < nks; ks++) {
            v8bf as[2][3];
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int p = 0; p < 3; p++) as[h][p] = an[h][p];
            const __bf16 *cur = (const __bf16 *)(bbuf + (size_t)(ks % 3) * kBufBytes);
            const __bf16 *nxt = (const __bf16 *)(bbuf + (size_t)((ks + 1) % 3) * kBufBytes);
            const __bf16 *cblk = Bnt + (size_t)min(ks + 2, nks - 1) * kMx2StepElems;
            unsigned char *cbuf = bbuf + (size_t)((ks + 2) % 3) * kBufBytes;
            const size_t ka = (size_t)min(ks + 1, nks - 1);
#pragma unroll
            for (int grp = 0; grp < 6; grp++) {
                // fragments two groups ahead (groups 0 and 1 of the next
                // K-step from the next buffer, which the barrier after group
                // 1 made complete)
                if (grp + 2 < 6)
                    read_group(bf[(grp + 2) % 3], cur, grp_pb(grp + 2), grp_h(grp + 2), q, c);
                else
                    read_group(bf[(grp + 2) % 3], nxt, grp_pb(grp - 4), grp_h(grp - 4), q, c);
                if (grp < 2) {
#pragma unroll
                    for (int p = 0; p < 3; p++)
                        an[grp][p] = *(const v8bf *)(Aw + (ka * 4 + 2 * grp) * 24 + p * 8);
                } else {
                    // pieces 0-2, 3-5, 6-8, 9-10 in groups 2..5
#pragma unroll
                    for (int i = 3 * (grp - 2); i < 3 * (grp - 2) + 3; i++)
                        if (i < kPiecesPerWave) stage_piece(cblk, cbuf, wave, lane, i);
                }
                __builtin_amdgcn_sched_barrier(0);
                const int h = grp_h(grp), pb = grp_pb(grp);
#pragma unroll
                for (int pa = 2 - pb; pa >= 0; pa--)
#pragma unroll
                    for (int m = 0; m < 7; m++)
                        acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as[h][pa], bf[grp % 3][m],
                                                                         acc[m], 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
                if (grp == 1) {
                    // the copy of block ks+1 (issued during K-step ks-1) has
                    // landed for this wavefront; after the barrier, for all
                    __builtin_amdgcn_s_waitcnt(kWaitAllButA);
                    __builtin_amdgcn_s_barrier();
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            // the next K-step's A fragments (issued before this K-step's
            // copy) have landed; the copy may still be in flight.  The
            // scheduling barriers keep the next K-step's register copies of
            // those fragments behind this wait (ahead of it the compiler
            // would wait for every outstanding load, the copy included).
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_waitcnt(kWaitAllButBlock);
            __builtin_amdgcn_sched_barrier(0);
        }
        // no copy may still be writing a buffer that the next N-tile reuses
        __builtin_amdgcn_s_waitcnt(kWaitAll);
        __builtin_amdgcn_s_barrier();

        const int col = nt * kNu + c;
        const float R[3] = {a.rf32[col], a.rf32[(size_t)a.ncol_pad + col],
                            a.rf32[2 * (size_t)a.ncol_pad + col]};
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const int rl = wave * 32 + 8 * (e >> 2) + 4 * q + (e & 3);
            const size_t p = row0 + rl;
            const bool valid = p < (size_t)a.npts;
            const float rp[3] = {valid ? (float)a.pts[3 * p] : 0.f,
                                 valid ? (float)a.pts[3 * p + 1] : 0.f,
                                 valid ? (float)a.pts[3 * p + 2] : 0.f};
            const __bf16 *pe = a.as3 + p * arow + (col >> 3) * 24 + (col & 7);
            const float ph = ((float)pe[0] + (float)pe[8]) + (float)pe[16];
            const size_t ei = p * lda + col;
            const float G[3] = {a.g32[ei], a.g32[gplane + ei], a.g32[2 * gplane + ei]};
            float v[32];
            const float t0 = acc[0][e] * ph;
            v[0] = t0;
#pragma unroll
            for (int d = 0; d < 3; d++) v[1 + d] = t0 * (rp[d] - R[d]);
#pragma unroll
            for (int bb = 0; bb < 3; bb++)
#pragma unroll
                for (int m = 0; m < 3; m++) v[4 + m + 3 * bb] = acc[1 + bb][e] * G[m];
#pragma unroll
            for (int d = 0; d < 3; d++)
#pragma unroll
                for (int m = 0; m < 3; m++) v[13 + m + 3 * d] = G[m] * acc[4 + d][e];
#pragma unroll
            for (int i = kNQ; i < 32; i++) v[i] = 0.f;
            const float sum = butterfly32(v, c);
            if (c < kNQ) qs[rl * kNQ + c] += (double)sum;
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    __syncthreads();
    double *out = a.part + ((size_t)set * a.nsplit + split) * kNQ * rows + row0;
    for (int i = tid; i < kRows * kNQ; i += 256) {
        const int qq = i / kRows, rl = i % kRows;
        out[(size_t)qq * rows + rl] = qs[rl * kNQ + qq];
    }
}

template <int NBUF>
static hipError_t launch_mx3(const ContractArgs &a, hipStream_t s) {
    static bool attr = false;
    const size_t lds = NBUF * kBufBytes + kLdsQ;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)contract_mx3_kernel<NBUF>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr = true;
    }
    if (a.rows % kRows || a.lda % 32 || a.ncol_pad != a.lda || a.ntiles * kNu != a.ncol_pad)
        return hipErrorInvalidValue;
    const long long groups = (a.rows / kRows + 7) / 8;
    const long long nwg = 8 * groups * a.nsplit * a.nmatset;
    if (nwg > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((contract_mx3_kernel<NBUF>), dim3((unsigned)nwg), dim3(256), lds, s, a);
    return hipGetLastError();
}

static hipError_t launch_mx3p(const ContractArgs &a, hipStream_t s) {
    static bool attr = false;
    const size_t lds = 3 * kBufBytes + kLdsQ;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)contract_mx3p_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr = true;
    }
    if (a.rows % kRows || a.lda % 32 || a.ncol_pad != a.lda || a.ntiles * kNu != a.ncol_pad)
        return hipErrorInvalidValue;
    const long long groups = (a.rows / kRows + 7) / 8;
    const long long nwg = 8 * groups * a.nsplit * a.nmatset;
    if (nwg > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(contract_mx3p_kernel, dim3((unsigned)nwg), dim3(256), lds, s, a);
    return hipGetLastError();
}

// nbuf 2: double-buffered; 3: the pipelined three-buffer form
hipError_t launch_contract_mx3(const ContractArgs &a, int nbuf, hipStream_t s) {
    if (nbuf == 2) return launch_mx3<2>(a, s);
    if (nbuf == 3) return launch_mx3p(a, s);
    return hipErrorInvalidValue;
}

}  // namespace gimic
