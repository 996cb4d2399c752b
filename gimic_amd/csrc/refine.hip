// Selection of the points that lie close to a nucleus (gfx950).
//
// The mixed-precision path evaluates J with FP32 products; next to a nucleus
// the tight basis functions make J large and its terms cancel, and FP32
// rounding there is the largest part of a plane flux's error.  Those points
// are evaluated again in FP64.  This file picks them: a deterministic stream
// compaction (per-block counts, one exclusive scan, in-order scatter), so the
// subset, its order, and every result bit repeat run to run.
#include "device_common.h"

namespace gimic {

namespace {

constexpr int kSelBlock = 256;

__device__ __forceinline__ bool near_any(const double *r, long long i, const double *coords,
                                         int natoms, double r2max) {
    const double x = r[3 * i], y = r[3 * i + 1], z = r[3 * i + 2];
    for (int a = 0; a < natoms; a++) {
        const double dx = x - coords[3 * a], dy = y - coords[3 * a + 1], dz = z - coords[3 * a + 2];
        if (dx * dx + dy * dy + dz * dz <= r2max) return true;
    }
    return false;
}

// counts[b] = number of selected points in block b (kSelBlock points)
__global__ __launch_bounds__(kSelBlock) void near_count_kernel(const double *r, long long npts,
                                                               const double *coords, int natoms,
                                                               double r2max, int *counts) {
    __shared__ int wsum[kSelBlock / 64];
    const long long i = (long long)blockIdx.x * kSelBlock + threadIdx.x;
    const bool f = i < npts && near_any(r, i, coords, natoms, r2max);
    const unsigned long long m = __ballot(f);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wsum[wave] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kSelBlock / 64; w++) t += wsum[w];
        counts[blockIdx.x] = t;
    }
}

// exclusive scan of counts[0..nb) into offs (one workgroup), total in *count
__global__ __launch_bounds__(1024) void near_scan_kernel(const int *counts, int nb, int *offs,
                                                         long long *count) {
    __shared__ long long part[1024];
    const int t = threadIdx.x;
    // each thread owns a contiguous range of blocks
    const int per = (nb + 1023) / 1024;
    const int b0 = min(nb, t * per), b1 = min(nb, b0 + per);
    long long s = 0;
    for (int b = b0; b < b1; b++) s += counts[b];
    part[t] = s;
    __syncthreads();
    for (int w = 1; w < 1024; w <<= 1) {
        const long long v = t >= w ? part[t - w] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long base = part[t] - s;  // exclusive prefix of this thread's range
    for (int b = b0; b < b1; b++) {
        offs[b] = (int)base;
        base += counts[b];
    }
    if (t == 1023) *count = part[1023];
}

__global__ __launch_bounds__(kSelBlock) void near_scatter_kernel(const double *r, long long npts,
                                                                 const double *coords, int natoms,
                                                                 double r2max, const int *offs,
                                                                 long long *idx) {
    __shared__ int wsum[kSelBlock / 64];
    const long long i = (long long)blockIdx.x * kSelBlock + threadIdx.x;
    const bool f = i < npts && near_any(r, i, coords, natoms, r2max);
    const unsigned long long m = __ballot(f);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wsum[wave] = __popcll(m);
    __syncthreads();
    int pos = offs[blockIdx.x] + __popcll(m & ((1ull << lane) - 1ull));
    for (int w = 0; w < wave; w++) pos += wsum[w];
    if (f) idx[pos] = i;
}

__global__ __launch_bounds__(256) void gather_idx_kernel(const double *r, const long long *idx,
                                                         long long n, double *out) {
    const long long j = (long long)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const long long i = idx[j];
#pragma unroll
    for (int d = 0; d < 3; d++) out[3 * j + d] = r[3 * i + d];
}

}  // namespace

size_t near_select_scratch(long long npts) {
    const long long nb = (npts + kSelBlock - 1) / kSelBlock;
    return (size_t)(2 * nb + 2);
}

hipError_t launch_near_select(const double *r, long long npts, const double *coords, int natoms,
                              double radius, long long *idx, long long *count, int *scratch,
                              hipStream_t s) {
    const long long nb = (npts + kSelBlock - 1) / kSelBlock;
    if (npts <= 0 || nb > 0x7fffffffLL) return hipErrorInvalidValue;
    // block offsets are int: at most 2^31 - 1 points
    if (npts > 0x7fffffffLL) return hipErrorInvalidValue;
    int *counts = scratch, *offs = scratch + nb;
    const double r2max = radius * radius;
    hipLaunchKernelGGL(near_count_kernel, dim3((unsigned)nb), dim3(kSelBlock), 0, s, r, npts,
                       coords, natoms, r2max, counts);
    hipLaunchKernelGGL(near_scan_kernel, dim3(1), dim3(1024), 0, s, counts, (int)nb, offs, count);
    hipLaunchKernelGGL(near_scatter_kernel, dim3((unsigned)nb), dim3(kSelBlock), 0, s, r, npts,
                       coords, natoms, r2max, offs, idx);
    return hipGetLastError();
}

hipError_t launch_gather_points(const double *r, const long long *idx, long long n, double *out,
                                hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_idx_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, r,
                       idx, n, out);
    return hipGetLastError();
}

}  // namespace gimic
