// Mixed-precision contraction with pre-split operands (gfx950).
//
// Arithmetic: an FP32 operand value is the exact sum of three BF16 parts
// h + m + l; the basis values keep kAParts of them and the right operand
// kBParts (device_common.h), and a block product keeps the part products of
// split order <= 2 (three and three: hh, hm, mh, hl, mm, lh), smallest
// first, on v_mfma_f32_16x16x32_bf16 with FP32 accumulation; FP64 partial
// sums are kept across N-tiles.  Both operands arrive split:
//   * the basis kernel writes the left operand as [row][k/8][kAParts][8]
//     (16 kAParts contiguous bytes per lane and K-step), and
//   * the right operand (constant per density) is split once on the host into
//     pair blocks of kBParts x 896 bytes per (density set, 8-function
//     K-group, 8-function N-group), [part][matrix][n][8 k]; a K-step's LDS
//     buffer gathers the 16 pair blocks of its four K-groups and the N-tile's
//     four N-groups.
// The K-loop therefore has no conversion work: each K-step streams the 16
// pair blocks of a later step into a free LDS buffer with gfx950's direct
// global-to-LDS loads (global_load_lds_dwordx4: no VGPRs, no ds_write; each
// lane names its own 16-byte source, so the gather costs nothing), loads
// the next A fragments, and runs KRB x 14 x split_products() MFMAs per
// wavefront against B fragments that are read from LDS one column block
// ahead.  NBUF = 2 streams block ks+1 during K-step ks; NBUF = 3 streams
// block ks+2, so a block has two K-steps to arrive from the Infinity Cache.
//
// Tile: 16 * KRB * WAVES point rows per workgroup (WAVES wavefronts x 16 KRB
// rows), 224 columns per N-tile (7 matrices x four 8-function groups), KRB x
// 14 MFMA blocks per wavefront.  KRB = 4 reads every B fragment from LDS once for 4
// row blocks (half the LDS traffic per MFMA of KRB = 2) and doubles the work
// per K-step that hides the next step's loads; it needs 224 accumulator
// registers.  WAVES = 8 (256 rows, two wavefronts per SIMD) lets one
// wavefront's MFMA chains run while the other waits on a dependency or a
// load, and each staged block feeds twice the rows.
//
// Per-point partial sums.  After the butterfly over the 16 columns of a lane
// group every lane holds all kNQ sums of its row; lane c writes sums c and
// c + 16 (FP64) to the row-major partial buffer [set][split][row][kNQ],
// 16 lanes of a group covering 128 contiguous bytes: an atomic store for
// the workgroup's first N-tile, no-return atomic adds for the later ones
// (one thread's atomics on one address keep program order), zeros from a
// workgroup without N-tiles.  Each address belongs to one lane of one
// workgroup, which adds its N-tiles in order, so the sums are the same, bit
// for bit, on every run, no clearing pass is needed, and no LDS holds them
// (three K-step buffers fit with 256-row tiles).

//
// Screening.  A workgroup visits only the live 8-function groups of its row
// tile (tiles.hip: the groups of the listed 32-function blocks in which
// screening leaves some function of the tile non-zero), on both sides of
// the product: an N-tile is four of them (their 32 functions x 7 matrices),
// the N-tiles are split evenly over the nsplit workgroups of the row tile,
// and the K-loop walks the same groups, four per K-step (an MFMA's K of
// 32).  On the 60-atom cage that is a fifth fewer (N-tile, K-step) pairs
// than N-tiles of whole 32-function blocks.  Each lane group of a wavefront
// reads its own group's A fragment (the group's place in the compressed
// row), and the buffer copy gathers the 16 pair blocks.  The LDS buffers
// rotate over K-steps, so the pipeline is the same as for a dense walk.
//
// Workgroup order.  Every workgroup re-reads its rows of A once per N-tile,
// and every row tile streams the B blocks of its listed (N, K) pairs.  The
// grid is 1-D and decoded so that (1) the nsplit workgroups that share a row
// tile (one per N-range) are dispatched back to back on the same XCD: the
// hardware hands workgroup i to XCD i mod 8, so slot s = i / 8 of XCD x runs
// the (s / nsplit)-th row tile of that XCD with split s % nsplit, and they
// read the same A rows at the same time, all but one from the XCD's L2; and
// (2) the row tiles go to the XCDs in turn (runs of kTileRun = 1), so an
// XCD runs tiles 8 apart along the points' Morton order: close enough that
// their lists mostly agree, and spread enough that its workgroups do not
// all stream the same B groups at once.  On the 60-atom cage, runs of 2, 4
// and 32 tiles were 0.3%, 1.3% and 12% slower; runs of 8 were 1-5% slower
// (two boxes).
#include "device_common.h"

namespace gimic {

namespace {

constexpr int kNu = 32;
constexpr int kCols = kMx2Cols;             // 224
constexpr int kCB = kCols / 16;             // 14 column blocks
constexpr int kBlk = kMx2BlockElems;        // BF16 per block / LDS buffer
constexpr size_t kBufBytes = (size_t)kBlk * 2;
// A K-step's LDS buffer is [group slot j][part][column][8]: slot j holds one
// K-group (kBParts x 3584 bytes), copied in kPPG 1 KiB pieces per slot, the
// last one partial when the slot does not fill it (three parts: 10752
// bytes, eleven pieces, the last for lanes 0-31).  Lane l of piece t of a
// slot fills unit u = 64 t + l of the slot: [part p][column], column =
// matrix m * 32 + 8 * N-slot + n.  As 224 = 7 * 32, the 32-unit chunk
// w = u / 32 = 2 t + l / 32 is p * 7 + m, and column % 32 = l % 32, so the
// lane's N-slot (l / 8) % 4 and function n = l % 8 are the same in every
// piece: the unit is element (p * 7 + m) * 64 + 8 n of the pair block of the
// slot's K-group and the lane's N-group, at lane offset nlo + 128 t
constexpr int kGroupBytes = kMx2GroupElems * 2;
constexpr int kPPG = (kGroupBytes + 1023) / 1024;           // pieces per group slot
constexpr int kLastLanes = (kGroupBytes - (kPPG - 1) * 1024) / 16;  // lanes of its last piece
constexpr int kPieces = 4 * kPPG;                            // pieces per K-step
constexpr int kPartElems = kCols * 8;                       // BF16 per part of a group
static_assert(4 * kGroupBytes <= kBlk * 2 && kGroupBytes % 16 == 0, "piece layout");

// pieces a wavefront copies of each block: wavefront w copies w, w + WAVES, ...
template <int WAVES>
constexpr int pieces_max() { return (kPieces + WAVES - 1) / WAVES; }
__device__ __forceinline__ int pieces_of(int wave, int waves) {
    return (kPieces - wave + waves - 1) / waves;
}

typedef __attribute__((address_space(3))) void lds_void;

__device__ __forceinline__ int pick4(int4 v, int i) {
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}

// the lane's offset, within a K-group's row of pair blocks, of its unit in
// piece 0 of a slot (fixed for an N-tile): pair block gn of its N-slot's
// group (ngn, a zero block, for padding), chunk l / 32, function l % 8
__device__ __forceinline__ int lane_offset(int4 en, int ngn, int lane) {
    const int x = pick4(en, (lane >> 3) & 3);
    const int gn = x < 0 ? ngn : x >> 12;
    return gn * kMx2PairElems + (lane >> 5) * 64 + (lane & 7) * 8;
}

// piece i of this wavefront's share of a K-step's buffer (64 lanes x 16
// bytes per instruction, LDS base in M0): piece q = kPPG j + t fills bytes
// [1024 t, 1024 t + 1024) of slot j from the row of K-group ek[j] >> 12
// (ek[j] < 0: the zero row nkg), each lane from its offset nlo + 128 t
template <int WAVES>
__device__ __forceinline__ void stage_piece_k8(const __bf16 *Bset, size_t rowst, int nkg, int4 ek,
                                               int nlo, unsigned char *buf, int wave, int lane,
                                               int i) {
    const int piece = wave + WAVES * i;
    const int j = piece / kPPG, t = piece - kPPG * j;
    const int v = pick4(ek, j);
    const __bf16 *src = Bset + (size_t)(v < 0 ? nkg : v >> 12) * rowst + 128 * t + nlo;
    if (t < kPPG - 1 || lane < kLastLanes)
        __builtin_amdgcn_global_load_lds((const void *)src,
                                         (lds_void *)(buf + j * kGroupBytes + t * 1024), 16, 0, 0);
}

template <int WAVES>
__device__ __forceinline__ void stage_block_k8(const __bf16 *Bset, size_t rowst, int nkg, int4 ek,
                                               int nlo, unsigned char *buf, int wave, int lane) {
    const int np = pieces_of(wave, WAVES);
#pragma unroll
    for (int i = 0; i < pieces_max<WAVES>(); i++)
        if (i < np) stage_piece_k8<WAVES>(Bset, rowst, nkg, ek, nlo, buf, wave, lane, i);
}

// One step of a transposing sum over the 16 lanes of a DPP row: lanes whose
// `up` bit is set keep the upper W of the values they still hold, the others
// the lower W, and each adds the copy its partner (DPP control CTRL, a lane
// that differs in that bit and holds the same values) does not keep.
template <int W, int CTRL>
__device__ __forceinline__ void xsum_step(float (&v)[32], bool up) {
#pragma unroll
    for (int i = 0; i < W; i++) {
        const float lo = v[i], hi = v[i + W];
        const float send = up ? lo : hi, keep = up ? hi : lo;
        v[i] = keep + GIMIC_DPP(send, CTRL);
    }
}

// v[0..31] summed over the 16 lanes of each DPP row, four exchanges of
// 16, 8, 4 and 2 values (row mirror, half-row mirror, quad xor 2, quad xor
// 1): lane c is left with the sums of slots 2c and 2c + 1 in v[0], v[1]
__device__ __forceinline__ void row16_xsum(float (&v)[32], int c) {
    xsum_step<16, 0x140>(v, (c & 8) != 0);
    xsum_step<8, 0x141>(v, (c & 4) != 0);
    xsum_step<4, 0x4E>(v, (c & 2) != 0);
    xsum_step<2, 0xB1>(v, (c & 1) != 0);
}

// the part products of split order <= 2, smallest first (three parts each:
// lh, mm, hl, mh, hm, hh)
__device__ __forceinline__ v4f mfma_split(const v8bf (&a)[kAParts], const v8bf (&b)[kBParts],
                                          v4f acc) {
#pragma unroll
    for (int ord = 2; ord >= 0; ord--)
#pragma unroll
        for (int ia = kAParts - 1; ia >= 0; ia--) {
            const int ib = ord - ia;
            if (ib >= 0 && ib < kBParts)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ia], b[ib], acc, 0, 0, 0);
        }
    return acc;
}

}  // namespace

// Row tiles per XCD: runs of kTileRun consecutive tiles go to the XCDs in
// turn, enough runs to cover every tile
constexpr int kTileRun = 1;
__host__ __device__ constexpr int tile_groups(int ntiles) {
    return (ntiles + 8 * kTileRun - 1) / (8 * kTileRun) * kTileRun;
}

// s_waitcnt immediates (gfx9 encoding: vmcnt [3:0] and [15:14], expcnt
// [6:4], lgkmcnt [11:8]): everything, or all but the last n vector-memory
// operations (the block copy issued last in a K-step)
constexpr int kWaitAll = 0;
constexpr int wait_all_but(int n) { return (n & 15) | ((n >> 4) << 14) | (7 << 4); }

template <int NBUF, int KRB, int WAVES>
__global__ __launch_bounds__(64 * WAVES, 1) void contract_mx2_kernel(ContractArgs a) {
    constexpr int kRows = 16 * KRB * WAVES;  // point rows per workgroup
    constexpr int kRB = KRB;                 // row blocks per wavefront
    constexpr int kPW = pieces_max<WAVES>();
    static_assert(kPW < 64, "vmcnt field");
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned char *bbuf = smem;
    float *rps = (float *)(smem + NBUF * kBufBytes);  // [kRows][3] point coordinates (FP32)

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = lane >> 4, c = lane & 15;
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int groups = tile_groups(a.rows / kRows);  // row tiles per XCD
    const int per_set = groups * a.nsplit;
    const int set = slot / per_set;
    const int split = slot % a.nsplit;
    const int jt = (slot % per_set) / a.nsplit;
    const int rtile = (jt / kTileRun) * 8 * kTileRun + xcd * kTileRun + jt % kTileRun;
    if (rtile >= a.rows / kRows) return;  // padding of the last run
    const size_t row0 = (size_t)rtile * kRows;
    const int lda = a.lda;
    const size_t rows = (size_t)a.rows;
    // N-tiles and K-steps of four live 8-function groups: entry = (global
    // group << 12) | position of the group in this tile's compressed rows;
    // -1 pads the last four (zero right operand)
    const int *K8 = a.k8list + (size_t)rtile * a.nblk * 4;
    const int nq = (a.k8count[rtile] + 3) >> 2;
    // This tile's basis rows: dense [row][lda], or compressed to the live
    // groups (a.toff non-null: the basis kernel wrote only those, so A and
    // the gradients are addressed by list position)
    const bool cmp = a.toff != nullptr;
    const size_t wcols = cmp ? (size_t)8 * a.k8count[rtile] : (size_t)lda;  // functions per row
    const size_t arow = kAParts * wcols;                                     // BF16 per split row
    const __bf16 *Atile = cmp ? a.as3 + (size_t)a.toff[rtile] * kRows * (8 * kAParts)
                              : a.as3 + row0 * arow;
    const float *Gtile = cmp ? a.g32 + (size_t)a.toff[rtile] * kRows * 24 : a.g32 + row0 * wcols;
    const size_t gpl = cmp ? (size_t)kRows * wcols : (size_t)a.gplane;  // gradient plane stride
    const int np = pieces_of(wave, WAVES);  // pieces this wavefront copies per block
    // this workgroup's partial sums, row-major [row][kNQ]
    double *pout = a.part + ((size_t)set * a.nsplit + split) * kNQ * rows;
    // the epilogue's point coordinates, once per workgroup (visible after
    // the first N-tile's barrier, before any epilogue)
    for (int i = tid; i < kRows; i += 64 * WAVES) {
        const size_t p = row0 + i;
        const bool valid = p < (size_t)a.npts;
#pragma unroll
        for (int d = 0; d < 3; d++) rps[3 * i + d] = valid ? (float)a.pts[3 * p + d] : 0.f;
    }

    // lane (g, c) of row block rb: row wave*32 + rb*16 + c, k-group entry g of the K-step
    const __bf16 *Arow = Atile + (size_t)(wave * 16 * kRB + c) * arow;
    auto a_off = [&](int4 v) -> size_t {  // this lane's group of a K-step, in BF16 of a row
        const int x = g == 0 ? v.x : g == 1 ? v.y : g == 2 ? v.z : v.w;
        return x < 0 ? 0 : (size_t)(cmp ? (x & 4095) : (x >> 12)) * (8 * kAParts);
    };
    // pair blocks of this density set: rows of K-groups (nkg: the zero row)
    // of ngn + 1 N-groups (ngn: the zero block)
    const int nkg = lda / 8, ngn = a.ncol_pad / 8;
    const size_t rowst = (size_t)(ngn + 1) * kMx2PairElems;
    const __bf16 *Bset = a.bs3 + (size_t)set * mx2_set_elems(lda, a.ncol_pad);
    const int e0 = (int)((long long)split * nq / a.nsplit);
    const int e1 = (int)((long long)(split + 1) * nq / a.nsplit);

    for (int e = e0; e < e1; e++) {
        const int4 en = *(const int4 *)(K8 + 4 * e);  // this N-tile's four groups
        const int nlo = lane_offset(en, ngn, lane);
        v4f acc[kRB][kCB];
#pragma unroll
        for (int rb = 0; rb < kRB; rb++)
#pragma unroll
            for (int cb = 0; cb < kCB; cb++) acc[rb][cb] = v4f{0.f, 0.f, 0.f, 0.f};

        const int4 q0 = *(const int4 *)K8, q1 = *(const int4 *)(K8 + (nq > 1 ? 4 : 0));
        stage_block_k8<WAVES>(Bset, rowst, nkg, q0, nlo, bbuf, wave, lane);
        if (NBUF == 3 && nq > 1)
            stage_block_k8<WAVES>(Bset, rowst, nkg, q1, nlo, bbuf + kBufBytes, wave, lane);
        v8bf an[kRB][kAParts];
        {
            const size_t ao = a_off(q0);
#pragma unroll
            for (int rb = 0; rb < kRB; rb++)
#pragma unroll
                for (int p = 0; p < kAParts; p++)
                    an[rb][p] = *(const v8bf *)(Arow + rb * 16 * arow + ao + p * 8);
        }
        __builtin_amdgcn_s_waitcnt(kWaitAll);
        __syncthreads();

        // groups of the next step's A fragments and of the next step's copy,
        // read one K-step early (the end-of-step wait covers the scalar loads)
        int4 ent_a = q1;
        int4 ent_c = *(const int4 *)(K8 + 4 * min(NBUF - 1, nq - 1));
        for (int ks = 0; ks < nq; ks++) {
            const int4 ea = ent_a, ec = ent_c;
            ent_a = *(const int4 *)(K8 + 4 * min(ks + 2, nq - 1));
            ent_c = *(const int4 *)(K8 + 4 * min(ks + NBUF, nq - 1));
            v8bf as[kRB][kAParts];
#pragma unroll
            for (int rb = 0; rb < kRB; rb++)
#pragma unroll
                for (int p = 0; p < kAParts; p++) as[rb][p] = an[rb][p];
            // Block streamed during this K-step.  With NBUF = 3 the copy is
            // issued every K-step (the last two repeat the last listed block
            // into a buffer nobody reads again) so that the count of outstanding
            // loads is the same on every iteration.
            // Vector-memory instructions (the next K-step's A fragments, two
            // per column block from the first one on, and the block copy's
            // pieces) are spread over the column blocks, between MFMA groups:
            // issued back to back they stall the in-order instruction stream,
            // and the matrix pipe behind it, while the vector-memory pipe
            // works through them.  With NBUF = 3 the copy comes after all A
            // loads (column blocks 3 to 13), so that waiting for "all but the
            // last copy" covers the A loads.
            const bool more = ks + 1 < nq;
            constexpr int kALoads = kAParts * kRB;
            constexpr int kCopyFrom = NBUF == 3 ? kCB - kPW : 0;
            static_assert(NBUF != 3 || (kALoads + 1) / 2 <= kCopyFrom, "A loads must precede the copy");
            const int ahead = ks + NBUF - 1;
            const bool copy = NBUF == 3 || ahead < nq;
            const size_t aoff = a_off(ea);
            unsigned char *cbuf = bbuf + (size_t)(ahead % NBUF) * kBufBytes;
            const __bf16 *cur = (const __bf16 *)(bbuf + (size_t)(ks % NBUF) * kBufBytes);
            v8bf bf[2][kBParts];
            {
                const int off = g * kMx2GroupElems + c * 8;
#pragma unroll
                for (int p = 0; p < kBParts; p++)
                    bf[0][p] = *(const v8bf *)(cur + p * kPartElems + off);
            }
#pragma unroll
            for (int cb = 0; cb < kCB; cb++) {
                const int sl = cb & 1;
                if (cb + 1 < kCB) {
                    const int off = g * kMx2GroupElems + ((cb + 1) * 16 + c) * 8;
#pragma unroll
                    for (int p = 0; p < kBParts; p++)
                        bf[sl ^ 1][p] = *(const v8bf *)(cur + p * kPartElems + off);
                }
#pragma unroll
                for (int j = 2 * cb; j < 2 * cb + 2; j++)
                    if (more && j < kALoads)
                        an[j / kAParts][j % kAParts] = *(const v8bf *)(
                            Arow + (j / kAParts) * 16 * arow + aoff + (j % kAParts) * 8);
                if (copy && cb >= kCopyFrom && cb - kCopyFrom < np)
                    stage_piece_k8<WAVES>(Bset, rowst, nkg, ec, nlo, cbuf, wave, lane, cb - kCopyFrom);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int rb = 0; rb < kRB; rb++) acc[rb][cb] = mfma_split(as[rb], bf[sl], acc[rb][cb]);
                __builtin_amdgcn_sched_barrier(0);
            }
            // this wavefront's A loads and the copy of block ks+1 have landed
            // (with NBUF = 3 the copy of block ks+2 may still be in flight);
            // after the barrier every wavefront's copy of block ks+1 is
            // visible and nobody reads the buffer the next copy overwrites
            // (a bare s_barrier: __syncthreads would add a fence that waits
            // for every outstanding load, the early copy included)
            if (NBUF == 3) {
                // np is uniform per wavefront: kPW, or kPW - 1 when the
                // pieces do not split evenly over the wavefronts
                if (np == kPW) __builtin_amdgcn_s_waitcnt(wait_all_but(kPW));
                else __builtin_amdgcn_s_waitcnt(wait_all_but(kPW > 1 ? kPW - 1 : 0));
            } else {
                __builtin_amdgcn_s_waitcnt(kWaitAll);
            }
            __builtin_amdgcn_s_barrier();
        }
        // no copy may still be writing a buffer that the next N-tile's
        // first copies reuse
        if (NBUF == 3) {
            __builtin_amdgcn_s_waitcnt(kWaitAll);
            __builtin_amdgcn_s_barrier();
        }

        // epilogue: lane (g, c) holds Y[row 4g + r][col c] of every block;
        // column block cb = 2 * matrix + half covers function c & 7 of the
        // N-tile's group 2 * half + (c >> 3) (padding: a zero column).
        // The global loads (phi parts and gradients) of each row are issued
        // one row ahead, so their latency overlaps the previous row's work.
        float R[2][3];
        int acol[2];  // column of the lane's function in this tile's basis rows
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int x = pick4(en, 2 * h + (c >> 3));
            const int fn = x < 0 ? 0 : (x >> 12) * 8 + (c & 7);
            acol[h] = x < 0 ? 0 : (cmp ? (x & 4095) : (x >> 12)) * 8 + (c & 7);
#pragma unroll
            for (int d = 0; d < 3; d++) R[h][d] = a.rf32[(size_t)d * a.ncol_pad + fn];
        }
        constexpr int kIt = 4 * kRB;  // rows per lane: it = 4 rb + r
        auto row_of = [&](int it) { return wave * 16 * kRB + (it >> 2) * 16 + 4 * g + (it & 3); };
        __bf16 pc[2][kAParts], pn[2][kAParts];
        float Gc[2][3], Gn[2][3];
        auto load_row = [&](int it, __bf16 (&pv)[2][kAParts], float (&Gv)[2][3]) {
            const size_t rl = row_of(it);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const __bf16 *pe = Atile + rl * arow + (acol[h] >> 3) * (8 * kAParts) + (acol[h] & 7);
#pragma unroll
                for (int k = 0; k < kAParts; k++) pv[h][k] = pe[8 * k];
                const size_t gi = rl * wcols + acol[h];
#pragma unroll
                for (int d = 0; d < 3; d++) Gv[h][d] = Gtile[d * gpl + gi];
            }
        };
        load_row(0, pc, Gc);
#pragma unroll
        for (int it = 0; it < kIt; it++) {
            const int rb = it >> 2, r = it & 3;
            {
                if (it + 1 < kIt) load_row(it + 1, pn, Gn);
                const int rl = row_of(it);
                const float rp[3] = {rps[3 * rl], rps[3 * rl + 1], rps[3 * rl + 2]};
                float q[kNQ];
#pragma unroll
                for (int i = 0; i < kNQ; i++) q[i] = 0.f;
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    // the basis value from its parts (three: exactly the FP32 value)
                    float ph = (float)pc[h][0];
#pragma unroll
                    for (int k = 1; k < kAParts; k++) ph += (float)pc[h][k];
                    const float *G = Gc[h];
                    const float t0 = acc[rb][h][r] * ph;
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
                // slots 2k, 2k + 1 = sums k, k + 16: lane c ends with the row
                // sums c and c + 16, the two it adds to the partial buffer
                float v[32];
#pragma unroll
                for (int k = 0; k < 16; k++) {
                    v[2 * k] = q[k];
                    v[2 * k + 1] = k + 16 < kNQ ? q[k + 16] : 0.f;
                }
                row16_xsum(v, c);
                const float s0 = v[0], s1 = v[1];
                double *o = pout + (row0 + rl) * kNQ;
                if (e == e0) {
                    __hip_atomic_store(o + c, (double)s0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (c + 16 < kNQ)
                        __hip_atomic_store(o + c + 16, (double)s1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    unsafeAtomicAdd(o + c, (double)s0);
                    if (c + 16 < kNQ) unsafeAtomicAdd(o + c + 16, (double)s1);
                }
            }
#pragma unroll
            for (int h = 0; h < 2; h++) {
#pragma unroll
                for (int k = 0; k < kAParts; k++) pc[h][k] = pn[h][k];
#pragma unroll
                for (int k = 0; k < 3; k++) Gc[h][k] = Gn[h][k];
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // a workgroup without N-tiles still owns its rows' sums for this split
    if (e0 == e1)
        for (int i = tid; i < kRows * kNQ; i += 64 * WAVES) pout[row0 * kNQ + i] = 0.0;
}

template <int NBUF, int KRB, int WAVES>
static hipError_t launch_one(const ContractArgs &a, hipStream_t s) {
    static bool attr = false;
    constexpr int rows = 16 * KRB * WAVES;
    const size_t lds = NBUF * kBufBytes + (size_t)rows * 3 * sizeof(float);
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)contract_mx2_kernel<NBUF, KRB, WAVES>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr = true;
    }
    if (a.rows % rows || a.lda % 32 || a.ncol_pad != a.lda || a.ntiles * kNu != a.ncol_pad)
        return hipErrorInvalidValue;
    const long long groups = tile_groups(a.rows / rows);
    const long long nwg = 8 * groups * a.nsplit * a.nmatset;
    if (nwg > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((contract_mx2_kernel<NBUF, KRB, WAVES>), dim3((unsigned)nwg),
                       dim3(64 * WAVES), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_contract_mx2(const ContractArgs &a, int nbuf, int krb, int waves, hipStream_t s) {
    if (waves == 4) {
        if (nbuf == 2 && krb == 2) return launch_one<2, 2, 4>(a, s);
        if (nbuf == 3 && krb == 2) return launch_one<3, 2, 4>(a, s);
        if (nbuf == 2 && krb == 4) return launch_one<2, 4, 4>(a, s);
    } else if (waves == 8 && krb == 2) {
        if (nbuf == 2) return launch_one<2, 2, 8>(a, s);
        if (nbuf == 3) return launch_one<3, 2, 8>(a, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace gimic
