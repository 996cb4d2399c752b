// Spatial point tiles and their screening-driven block lists (gfx950).
//
// Basis functions are zero beyond their contraction's screening radius
// (basis.f90:118-136), so over a compact tile of points most 32-function
// blocks of the basis are exactly zero in both the values phi and the
// gradients.  A zero K-block adds exact zeros to Y = Phi * B, and a zero
// N-block is multiplied by zeros in the epilogue, so a contraction
// workgroup that walks only the blocks that can be non-zero over its tile
// produces the same numbers as the dense product.
//
// Two steps per chunk of points:
//  1. Order the points along a Morton curve of 32^3 cubic cells over their
//     bounding box (a stable LSD radix sort with 4-bit digits of the 15-bit
//     cell key, so the order - and with it every result bit - is the same
//     on every run); the basis, contraction and finalize kernels work on the
//     sorted rows, and finalize writes each tensor back to its caller index.
//  2. For every row tile of the contraction kernel, list in ascending order
//     the blocks holding a function whose centre lies within its screening
//     radius of the tile's bounding box (a conservative test: a block that
//     is zero over the tile may be listed, one that is not is never left
//     out).
#include "device_common.h"

namespace gimic {

namespace {

constexpr int kSortBlock = 1024;  // elements per radix block (4 rounds of 256)
constexpr int kBins = 16;         // 4-bit digits
constexpr int kCellBits = 5;      // 32 cells per axis

__device__ __forceinline__ unsigned spread3(unsigned v) {
    // bits b0..b4 -> positions 0, 3, 6, 9, 12
    unsigned r = 0;
#pragma unroll
    for (int i = 0; i < kCellBits; i++) r |= ((v >> i) & 1u) << (3 * i);
    return r;
}

// min/max of the points, one partial box per workgroup: out[blk][6]
__global__ __launch_bounds__(256) void bbox_partial_kernel(const double *pts, int n, double *out) {
    __shared__ double s[6][256];
    const int t = threadIdx.x;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (long long i = (long long)blockIdx.x * 256 + t; i < n; i += (long long)gridDim.x * 256)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double x = pts[3 * i + d];
            lo[d] = fmin(lo[d], x);
            hi[d] = fmax(hi[d], x);
        }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        s[d][t] = lo[d];
        s[3 + d][t] = hi[d];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                s[d][t] = fmin(s[d][t], s[d][t + w]);
                s[3 + d][t] = fmax(s[3 + d][t], s[3 + d][t + w]);
            }
        __syncthreads();
    }
    if (t < 6) out[blockIdx.x * 6 + t] = s[t][0];
}

// Morton key of each point's cell; keys[i] = key, idx[i] = i
__global__ __launch_bounds__(256) void cell_key_kernel(const double *pts, int n, const double *bpart,
                                                       int nbpart, unsigned *keys, int *idx) {
    __shared__ double s[6][256];
    const int t = threadIdx.x;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        s[d][t] = t < nbpart ? bpart[t * 6 + d] : INFINITY;
        s[3 + d][t] = t < nbpart ? bpart[t * 6 + 3 + d] : -INFINITY;
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                s[d][t] = fmin(s[d][t], s[d][t + w]);
                s[3 + d][t] = fmax(s[3 + d][t], s[3 + d][t + w]);
            }
        __syncthreads();
    }
    const double lo[3] = {s[0][0], s[1][0], s[2][0]};
    double h = fmax(fmax(s[3][0] - lo[0], s[4][0] - lo[1]), s[5][0] - lo[2]) / (1 << kCellBits);
    if (!(h > 0.0) || !isfinite(h)) h = 1.0;
    const long long i = (long long)blockIdx.x * 256 + t;
    if (i >= n) return;
    unsigned key = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        // fmax/fmin map NaN to the first cell
        const double c = fmin(fmax((pts[3 * i + d] - lo[d]) / h, 0.0), (double)((1 << kCellBits) - 1));
        key |= spread3((unsigned)c) << d;
    }
    keys[i] = key;
    idx[i] = (int)i;
}

// per radix block digit counts, digit-major: hist[d * nb + b]
__global__ __launch_bounds__(256) void radix_hist_kernel(const unsigned *keys, int n, int shift,
                                                         int *hist) {
    __shared__ int cnt[kBins];
    const int t = threadIdx.x, b = blockIdx.x, nb = gridDim.x;
    if (t < kBins) cnt[t] = 0;
    __syncthreads();
    for (int r = 0; r < kSortBlock / 256; r++) {
        const long long i = (long long)b * kSortBlock + r * 256 + t;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (kBins - 1)], 1);
    }
    __syncthreads();
    if (t < kBins) hist[t * nb + b] = cnt[t];
}

// exclusive scan of m counts in place (one workgroup of 1024): the counts
// are read in coalesced passes of 1024, each scanned within wavefronts
// (shuffles), then across the 16 wavefronts, plus the running total
__global__ __launch_bounds__(1024) void radix_scan_kernel(int *v, int m) {
    __shared__ int wsum[16];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int carry = 0;
    for (int base = 0; base < m; base += 1024) {
        const int i = base + t;
        const int x = i < m ? v[i] : 0;
        int inc = x;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(inc, d);
            if (lane >= d) inc += y;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        int before = carry, total = carry;
#pragma unroll
        for (int w = 0; w < 16; w++) {
            const int ws = wsum[w];
            if (w < wave) before += ws;
            total += ws;
        }
        if (i < m) v[i] = before + inc - x;
        carry = total;
        __syncthreads();  // wsum is rewritten by the next pass
    }
}

// stable scatter by digit: element order within a digit is kept
__global__ __launch_bounds__(256) void radix_scatter_kernel(const unsigned *kin, const int *iin, int n,
                                                            int shift, const int *off, unsigned *kout,
                                                            int *iout) {
    __shared__ int wcnt[4][kBins];
    __shared__ int run[kBins];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, b = blockIdx.x, nb = gridDim.x;
    if (t < kBins) run[t] = off[t * nb + b];
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int r = 0; r < kSortBlock / 256; r++) {
        const long long i = (long long)b * kSortBlock + r * 256 + t;
        const bool valid = i < n;
        const unsigned key = valid ? kin[i] : 0u;
        const int dig = (int)((key >> shift) & (kBins - 1));
        int rank = 0;
#pragma unroll
        for (int v = 0; v < kBins; v++) {
            const unsigned long long m = __ballot(valid && dig == v);
            if (dig == v) rank = __popcll(m & below);
            if (lane == 0) wcnt[wave][v] = __popcll(m);
        }
        __syncthreads();  // wcnt written, run up to date
        if (valid) {
            int pos = run[dig] + rank;
            for (int w = 0; w < wave; w++) pos += wcnt[w][dig];
            kout[pos] = key;
            iout[pos] = iin[i];
        }
        __syncthreads();  // every thread has read run and wcnt
        if (t < kBins) run[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
        __syncthreads();  // run updated before the next round's wcnt is written
    }
}

// sorted points and their caller indices
__global__ __launch_bounds__(256) void gather_points_kernel(const double *pts, const int *idx, int n,
                                                            double *spts, int *perm) {
    const long long j = (long long)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const int i = idx[j];
    perm[j] = i;
#pragma unroll
    for (int d = 0; d < 3; d++) spts[3 * j + d] = pts[3 * (long long)i + d];
}

// One workgroup per row tile: bounding box of its valid rows, then the
// ascending list of blocks with a function that may be non-zero there.
constexpr int kMaxBlk = 1024;  // 32-function blocks per row (32768 functions)

__global__ __launch_bounds__(256) void block_list_kernel(TileArgs a) {
    __shared__ double s[6][256];
    __shared__ int wsum[4];
    __shared__ unsigned char gmask[kMaxBlk];  // live 8-function groups of each block
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const long long r0 = (long long)blockIdx.x * a.tile_rows;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int r = t; r < a.tile_rows; r += 256) {
        const long long p = r0 + r;
        if (p < a.npts)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double x = a.spts[3 * p + d];
                lo[d] = fmin(lo[d], x);
                hi[d] = fmax(hi[d], x);
            }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        s[d][t] = lo[d];
        s[3 + d][t] = hi[d];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                s[d][t] = fmin(s[d][t], s[d][t + w]);
                s[3 + d][t] = fmax(s[3 + d][t], s[3 + d][t + w]);
            }
        __syncthreads();
    }
    const double blo[3] = {s[0][0], s[1][0], s[2][0]}, bhi[3] = {s[3][0], s[4][0], s[5][0]};
    const bool any = r0 < a.npts;
    int *list = a.blist + (long long)blockIdx.x * a.nblk;
    int base = 0;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int b0 = 0; b0 < a.nblk; b0 += 256) {
        const int b = b0 + t;
        bool act = false;
        unsigned m4 = 0;  // groups of 8 functions that a live shell touches
        if (any && b < a.nblk) {
            for (int k = a.blk_beg[b]; k < a.blk_beg[b + 1] && m4 != 15u; k++) {
                const DevShell sh = a.shells[a.blk_sh[k]];
                const double *R = a.coords + 3 * sh.atom;
                double d2 = 0.0;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const double e = fmax(0.0, fmax(blo[d] - R[d], R[d] - bhi[d]));
                    d2 += e * e;
                }
                // relative slack for rounding in the basis kernel's own distance
                if (sqrt(d2) <= sh.radius * (1.0 + 1e-12) + 1e-12) {
                    const int f0 = max(0, sh.fout - 32 * b);
                    const int f1 = min(31, sh.fout + sh.ncomp_out - 1 - 32 * b);
                    for (int gq = f0 >> 3; gq <= (f1 >> 3); gq++) m4 |= 1u << gq;
                }
            }
            act = m4 != 0u;
        }
        if (b < a.nblk) gmask[b] = (unsigned char)m4;
        const unsigned long long m = __ballot(act);
        if (lane == 0) wsum[wave] = __popcll(m);
        __syncthreads();
        int pos = base + __popcll(m & below);
        for (int w = 0; w < wave; w++) pos += wsum[w];
        if (act) list[pos] = b;
        base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    if (t == 0) a.bcount[blockIdx.x] = base;
    // the live groups in list order, four per K-step (the last padded with -1)
    __syncthreads();
    if (t == 0) {
        int *kl = a.k8list + (long long)blockIdx.x * a.nblk * 4;
        int k = 0;
        for (int b = 0; b < a.nblk; b++) {
            const unsigned m = gmask[b];
            if (!m) continue;
            for (int gq = 0; gq < 4; gq++)
                if (m >> gq & 1u) {
                    kl[k] = ((4 * b + gq) << 12) | k;
                    k++;
                }
        }
        a.k8count[blockIdx.x] = k;
        while (k & 3) kl[k++] = -1;
    }
}

// exclusive prefix sum of the tiles' live group counts, toff[0 .. n]
__global__ __launch_bounds__(256) void tile_offsets_kernel(const int *cnt, int n, int *toff) {
    __shared__ int s[256];
    const int t = threadIdx.x;
    const int per = (n + 255) / 256;
    const int beg = min(n, t * per), end = min(n, beg + per);
    int sum = 0;
    for (int i = beg; i < end; i++) sum += cnt[i];
    s[t] = sum;
    __syncthreads();
    for (int w = 1; w < 256; w <<= 1) {
        const int x = t >= w ? s[t - w] : 0;
        __syncthreads();
        s[t] += x;
        __syncthreads();
    }
    int run = s[t] - sum;
    for (int i = beg; i < end; i++) {
        toff[i] = run;
        run += cnt[i];
    }
    if (t == 255) toff[n] = s[255];
}

}  // namespace

hipError_t launch_tile_offsets(const int *cnt, int ntiles, int *toff, hipStream_t s) {
    if (ntiles <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(tile_offsets_kernel, dim3(1), dim3(256), 0, s, cnt, ntiles, toff);
    return hipGetLastError();
}

int sort_scratch_ints(int n) {
    const int nb = (n + kSortBlock - 1) / kSortBlock;
    return 4 * n + kBins * nb;  // two key/index ping-pong pairs + digit counts
}

hipError_t launch_sort_points(const double *pts, int n, double *spts, int *perm, double *bpart,
                              int *scratch, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int nb = (n + kSortBlock - 1) / kSortBlock;
    unsigned *k0 = (unsigned *)scratch, *k1 = k0 + n;
    int *i0 = scratch + 2 * n, *i1 = i0 + n, *hist = scratch + 4 * n;
    const int npart = n < 256 * 256 ? (n + 255) / 256 : 256;
    hipLaunchKernelGGL(bbox_partial_kernel, dim3(npart), dim3(256), 0, s, pts, n, bpart);
    hipLaunchKernelGGL(cell_key_kernel, dim3((n + 255) / 256), dim3(256), 0, s, pts, n,
                       (const double *)bpart, npart, k0, i0);
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 4 * pass;
        unsigned *kin = pass & 1 ? k1 : k0, *kout = pass & 1 ? k0 : k1;
        int *iin = pass & 1 ? i1 : i0, *iout = pass & 1 ? i0 : i1;
        hipLaunchKernelGGL(radix_hist_kernel, dim3(nb), dim3(256), 0, s, (const unsigned *)kin, n,
                           shift, hist);
        hipLaunchKernelGGL(radix_scan_kernel, dim3(1), dim3(1024), 0, s, hist, kBins * nb);
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(nb), dim3(256), 0, s, (const unsigned *)kin,
                           (const int *)iin, n, shift, (const int *)hist, kout, iout);
    }
    hipLaunchKernelGGL(gather_points_kernel, dim3((n + 255) / 256), dim3(256), 0, s, pts,
                       (const int *)i0, n, spts, perm);
    return hipGetLastError();
}

hipError_t launch_block_lists(const TileArgs &a, hipStream_t s) {
    if (a.tile_rows <= 0 || a.rows % a.tile_rows || a.nblk <= 0 || a.nblk > kMaxBlk)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(block_list_kernel, dim3(a.rows / a.tile_rows), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace gimic
