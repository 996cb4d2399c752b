// Plane-flux inputs built on the device (gfx950).
//
// plane_index_kernel gives every point of a batch of planes its plane
// (binary search in the plane offsets), so a caller's host point arrays need
// no per-point index array.  plane_grid_kernel generates the points and
// weights of bond-plane product grids themselves (grid.f90:478-495,
// gaussint.f90): point (i, j) of plane k is
//     ((o + p1_i v1) + p2_j v2) + p3 v3,   weight (w3 w2_j) w1_i,
// point index j * n1 + i within the plane (i fastest), rounded at every
// step exactly as the host grid code (Grid.points / Grid.weights), so the
// points - and every flux bit - equal those of the host-built grid.  A scan
// that moves its planes then sends only the per-plane descriptors and 1-D
// quadrature rules, and the whole scan runs as one graph.
#include "device_common.h"

namespace gimic {

namespace {

// last k with offs[k] <= i, for 0 <= i < offs[n]
__device__ __forceinline__ int plane_of(const long long *offs, int n, long long i) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (offs[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void plane_index_kernel(const long long *offs, int nplanes,
                                                          long long npts, int *plane) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < npts) plane[i] = plane_of(offs, nplanes, i);
}

__global__ __launch_bounds__(256) void plane_grid_kernel(PlaneGridArgs a, double *r, double *w,
                                                         int *plane) {
    // every product and sum rounded on its own, as the host grid code does
    // (hipcc contracts a * b + c into one FMA by default)
#pragma clang fp contract(off)
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= a.npts) return;
    const int k = plane_of(a.poff, a.nplanes, idx);
    const long long local = idx - a.poff[k];
    const int n1 = a.n1[k];
    const long long i = local % n1, j = local / n1;
    const double p1 = a.nodes1[a.off1[k] + i], p2 = a.nodes2[a.off2[k] + j];
    const double *o = a.origin + 3 * k, *ax = a.axes + 9 * k;
    const double p3 = a.p3w3[2 * k], w3 = a.p3w3[2 * k + 1];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double x = o[d] + p1 * ax[d];
        x = x + p2 * ax[3 + d];
        r[3 * idx + d] = x + p3 * ax[6 + d];
    }
    w[idx] = (w3 * a.wts2[a.off2[k] + j]) * a.wts1[a.off1[k] + i];
    plane[idx] = k;
}

}  // namespace

hipError_t launch_plane_index(const long long *offs, int nplanes, long long npts, int *plane,
                              hipStream_t s) {
    if (npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(plane_index_kernel, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s,
                       offs, nplanes, npts, plane);
    return hipGetLastError();
}

hipError_t launch_plane_grid(const PlaneGridArgs &a, double *r, double *w, int *plane,
                             hipStream_t s) {
    if (a.npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(plane_grid_kernel, dim3((unsigned)((a.npts + 255) / 256)), dim3(256), 0, s,
                       a, r, w, plane);
    return hipGetLastError();
}

}  // namespace gimic
