// Tensor assembly from the contraction partials, and plane-flux reductions.
//
// J_mb = 1/2 [A_mb + GIAO (eps_bcd r_c Q_md + eps_bmd C_d)] + 1/2 eps_mbc rho r_c
// reproduces jtensor.F90:163-235 (ZETA = 1/2, globals.f90:47): A is the
// paramagnetic term (ppd), the GIAO bracket is prsp1 + prsp2, and the last
// term is the diamagnetic dpd added to the off-diagonal elements.  The
// paramag switch zeroes the bracket, the diamag switch the last term.
// All of this runs in FP64 whatever the contraction precision.
//
// The flux contribution of a point is w_p n.(J B) (integral.f90:119-143),
// zero outside the integration radius around the plane centre; the signed
// modulus contribution is w_p sgn(n.jv) |jv| with jv = J B and a 1e-12 dead
// zone on the sign (integral.f90:258-283), and the ACID density w_p dT^2
// (acid.f90:11-40, integral.f90:414-511), all from the same tensor; per-plane
// sums (net, positive part, negative part) use a fixed-order tree so results
// are reproducible run to run.
#include "device_common.h"

namespace gimic {

__device__ __forceinline__ double eps3(int i, int j, int k) {
    return (double)((i - j) * (j - k) * (k - i)) * 0.5;
}

__global__ __launch_bounds__(256) void finalize_kernel(FinalizeArgs a) {
    // row-major partials: the workgroup's 256 rows are 256 kNQ contiguous
    // doubles per (set, split), staged through LDS with coalesced 16-byte
    // loads (a thread's own row is 176 bytes: read directly, every load of
    // a wavefront would touch 64 separate pieces)
    __shared__ double2 stage[256 * kNQ / 2];
    const int p = blockIdx.x * 256 + threadIdx.x;
    const bool valid = p < a.npts;
    double r[3] = {0.0, 0.0, 0.0};
    long long o = 0;
    if (valid) {
        r[0] = a.pts[3 * p];
        r[1] = a.pts[3 * p + 1];
        r[2] = a.pts[3 * p + 2];
        o = a.perm ? a.perm[p] : p;  // caller index of row p
        if (a.pidx) o = a.pidx[o];   // its place in the output arrays
    }
    for (int set = 0; set < a.nmatset; set++) {
        double q[kNQ];
#pragma unroll
        for (int i = 0; i < kNQ; i++) q[i] = 0.0;
        for (int s = 0; s < a.nsplit; s++) {
            const double *src = a.part + ((size_t)set * a.nsplit + s) * kNQ * (size_t)a.rows;
            if (a.part_rows) {
                const double2 *blk = (const double2 *)(src + (size_t)blockIdx.x * 256 * kNQ);
                __syncthreads();  // the previous (set, split) has been read out
                for (int i = threadIdx.x; i < 256 * kNQ / 2; i += 256) stage[i] = blk[i];
                __syncthreads();
#pragma unroll
                for (int j = 0; j < kNQ / 2; j++) {
                    const double2 t = stage[threadIdx.x * (kNQ / 2) + j];
                    q[2 * j] += t.x;
                    q[2 * j + 1] += t.y;
                }
            } else if (valid) {
#pragma unroll
                for (int i = 0; i < kNQ; i++) q[i] += src[(size_t)i * a.rows + p];
            }
        }
        if (!valid) continue;
        // q[1+c] = E_c = rho r_c - C_c, C_c = sum_nu Y0 phi R_nu,c.  With all
        // three terms on, eps_bmd C_d + eps_mbc rho r_c = eps_mbd E_d, so the
        // tensor is assembled without forming rho r and C separately.
        const double rho = q[0];
        const bool full = a.giao && a.paramag && a.diamag;
        double J[9];
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int m = 0; m < 3; m++) {
                double rq = 0.0;  // eps_bcd r_c Q_md
#pragma unroll
                for (int cc = 0; cc < 3; cc++)
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const double e = eps3(b, cc, d);
                        if (e != 0.0) rq += e * r[cc] * q[13 + m + 3 * d];
                    }
                double v;
                if (full) {
                    double ee = 0.0;
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const double e = eps3(m, b, d);
                        if (e != 0.0) ee += e * q[1 + d];
                    }
                    v = 0.5 * (q[4 + m + 3 * b] + rq) + 0.5 * ee;
                } else {
                    double para = q[4 + m + 3 * b];
                    if (a.giao) {
                        para += rq;
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            const double e = eps3(b, m, d);
                            if (e != 0.0) para += e * (rho * r[d] - q[1 + d]);
                        }
                    }
                    v = a.paramag ? 0.5 * para : 0.0;
                    if (a.diamag) {
#pragma unroll
                        for (int cc = 0; cc < 3; cc++) {
                            const double e = eps3(m, b, cc);
                            if (e != 0.0) v += 0.5 * e * rho * r[cc];
                        }
                    }
                }
                J[m + 3 * b] = v;
            }
        double *out = a.jt + (size_t)set * a.jt_stride + (size_t)o * 9;
#pragma unroll
        for (int i = 0; i < 9; i++) out[i] = J[i];
        if (set == 0 && a.contrib) {
            const int pl = a.plane ? a.plane[o] : 0;
            const double *n = a.pnorm + 3 * pl, *B = a.pbfield + 3 * pl, *ctr = a.pcenter + 3 * pl;
            double dot = 0.0;
#pragma unroll
            for (int m = 0; m < 3; m++)
                dot += n[m] * (J[m] * B[0] + J[m + 3] * B[1] + J[m + 6] * B[2]);
            const double dx = r[0] - ctr[0], dy = r[1] - ctr[1], dz = r[2] - ctr[2];
            const double dist = sqrt(dx * dx + dy * dy + dz * dz);
            const bool outside = dist > a.pradius[pl];
            a.contrib[o] = outside ? 0.0 : dot * a.w[o];
            if (a.contrib_mod) {
                double jv2 = 0.0;
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    const double v = J[m] * B[0] + J[m + 3] * B[1] + J[m + 6] * B[2];
                    jv2 += v * v;
                }
                const double sgn = fabs(dot) < 1e-12 ? 0.0 : (dot > 0.0 ? 1.0 : -1.0);
                a.contrib_mod[o] = outside ? 0.0 : sgn * sqrt(jv2) * a.w[o];
            }
            if (a.contrib_acid) {
                // get_acid, acid.f90:11-40 (DP33 = 0.3333333 as in the reference)
                const double d1 = J[0] - J[4], d2 = J[4] - J[8], d3 = J[8] - J[0];
                const double o1 = J[3] + J[1], o2 = J[6] + J[2], o3 = J[7] + J[5];
                const double dt2 = 0.3333333 * (d1 * d1 + d2 * d2 + d3 * d3) +
                                   0.5 * (o1 * o1 + o2 * o2 + o3 * o3);
                a.contrib_acid[o] = outside ? 0.0 : dt2 * a.w[o];
            }
        }
    }
}

__global__ __launch_bounds__(256) void plane_reduce_kernel(const double *contrib,
                                                           const long long *offs, double *out) {
    __shared__ double s[3][256];
    const int pl = blockIdx.x, t = threadIdx.x;
    const long long beg = offs[pl], end = offs[pl + 1];
    double net = 0.0, pos = 0.0, neg = 0.0;
    for (long long i = beg + t; i < end; i += 256) {
        const double v = contrib[i];
        net += v;
        if (v > 0.0) pos += v;
        else neg += v;
    }
    s[0][t] = net;
    s[1][t] = pos;
    s[2][t] = neg;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            s[0][t] += s[0][t + w];
            s[1][t] += s[1][t + w];
            s[2][t] += s[2][t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        out[3 * pl] = s[0][0];
        out[3 * pl + 1] = s[1][0];
        out[3 * pl + 2] = s[2][0];
    }
}

// Shielding and magnetizability quadrature on an external grid
// (jfield.f90:584-929).  Workgroup (p, k): point partition p (the points of
// one atom's grid, nelpts.info), and nucleus k < nnuc (shielding) or
// k == nnuc (magnetizability).  With j_b = J (-e_b) the field-direction
// current (column b of J negated), d = r - R_k and
// f = -1 / (|d|^3 c^2), c = 137.0359998 (jfield.f90:706):
//   shielding:       s_b = 1e6 f (d x j_b)_b        (:712-729)
//   magnetizability: s_b = 1/2 (r x j_b)_b          (:806-822)
// Per partition: sum w s_x, sum w s_y, sum w s_z, and the positive and
// negative parts of sum w (s_x + s_y + s_z) (:731-736, :824-829), summed with
// a fixed-order tree so results are the same on every run.
__global__ __launch_bounds__(256) void property_kernel(const double *jt, const double *r,
                                                       const double *w, const long long *offs,
                                                       const double *nuc, int nnuc, int npart,
                                                       double *out) {
    __shared__ double sm[5][256];
    const int part = blockIdx.x, k = blockIdx.y, t = threadIdx.x;
    const bool shield = k < nnuc;
    const double c2 = 137.0359998 * 137.0359998;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (long long i = offs[part] + t; i < offs[part + 1]; i += 256) {
        const double *J = jt + 9 * i;
        double d[3] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
        double f = 0.5;
        if (shield) {
#pragma unroll
            for (int e = 0; e < 3; e++) d[e] -= nuc[3 * k + e];
            const double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            f = 1.0e6 * (-1.0 / (d2 * sqrt(d2)) / c2);
        }
        double s[3];
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const int b1 = (b + 1) % 3, b2 = (b + 2) % 3;
            // j = J (-e_b): j_m = -J[m + 3b]
            s[b] = f * (d[b1] * -J[b2 + 3 * b] - d[b2] * -J[b1 + 3 * b]);
        }
        const double pd = s[0] + s[1] + s[2];
#pragma unroll
        for (int b = 0; b < 3; b++) acc[b] += w[i] * s[b];
        if (pd >= 0.0) acc[3] += pd * w[i];
        else acc[4] += pd * w[i];
    }
#pragma unroll
    for (int q = 0; q < 5; q++) sm[q][t] = acc[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h)
#pragma unroll
            for (int q = 0; q < 5; q++) sm[q][t] += sm[q][t + h];
        __syncthreads();
    }
    if (t < 5) out[((size_t)k * npart + part) * 5 + t] = sm[t][0];
}

hipError_t launch_properties(const double *jt, const double *r, const double *w,
                             const long long *offs, const double *nuc, int nnuc, int npart,
                             double *out, hipStream_t s) {
    if (npart <= 0) return hipSuccess;
    hipLaunchKernelGGL(property_kernel, dim3(npart, nnuc + 1), dim3(256), 0, s, jt, r, w, offs, nuc,
                       nnuc, npart, out);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void fill_zero_kernel(double *p, long long n) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        p[i] = 0.0;
}

hipError_t launch_fill_zero(double *p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, (long long)n);
    return hipGetLastError();
}

hipError_t launch_finalize(const FinalizeArgs &a, hipStream_t s) {
    if (a.npts <= 0) return hipSuccess;
    // row-major partials are read in whole workgroups of 256 rows
    if (a.part_rows && (a.rows % 256 || a.npts > a.rows)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(finalize_kernel, dim3((a.npts + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_plane_reduce(const double *contrib, const long long *offs, int nplanes,
                               double *out, hipStream_t s) {
    if (nplanes <= 0) return hipSuccess;
    hipLaunchKernelGGL(plane_reduce_kernel, dim3(nplanes), dim3(256), 0, s, contrib, offs, out);
    return hipGetLastError();
}

}  // namespace gimic
