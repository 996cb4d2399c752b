// Basis-function evaluation per point tile (gfx950).
//
// One thread per (point, contraction), contraction index fastest, so a
// wavefront writes neighbouring function columns of one point row.  Each
// thread evaluates the contracted radial parts g = sum c e^{-a s} and
// h = sum a c e^{-a s} once (the reference recomputes the exponentials for
// every derivative axis, caos.f90:39-64) and forms, per cartesian component
// x^a y^b z^c (gtodefs.f90:86-123):
//   phi    = x^a y^b z^c g                                (caos.f90:17-37)
//   d_m phi = k_m r_m^(k_m-1)(...) g - 2 r_m x^a y^b z^c h (caos.f90:39-64)
// with contractions outside their screening radius set to zero
// (basis.f90:118-136).  Spherical contractions are T_l times the cartesian
// ones.  Outputs: phi (the contraction's left operand) and its gradient in
// FP64 for the FP64 path.
//
// The split output of the mixed kernels (phi as three BF16 parts plus FP32
// gradients), and the FP64 rows of bases up to 4,800 functions, are
// produced by basis_eval_staged_kernel: a workgroup
// owns a tile of 1-32 whole point rows, its threads evaluate (point,
// contraction) pairs into a zeroed LDS image of those rows, and the rows
// then leave with 16-byte coalesced stores.  Written directly, every store
// instruction of a wavefront scatters 64 short values over a whole row.
#include "device_common.h"

namespace gimic {

__constant__ double c_sph[466];  // T_l, l = 2..5, (2l+1) x ncart(l) row-major

// Radial parts g = sum c e^{-a s} and h = sum a c e^{-a s} of one
// contraction.  The primitives' exponents and coefficients are loaded in
// groups of four before any exponential, so a group costs one memory
// latency rather than one per primitive (the terms are summed in the same
// order as a plain loop).
// e^x to FP32 accuracy (relative error ~1e-7) for the mixed path, whose
// basis values are rounded to FP32 anyway: x log2(e) = n + f in FP64 with n
// integer and |f| <= 1/2, 2^f from the hardware exp2 (v_exp_f32), scaled by
// 2^n in FP64 (no FP32 overflow or underflow for any n)
__device__ __forceinline__ double exp_f32acc(double x) {
    const double t = x * 1.4426950408889634;  // log2(e)
    const double n = rint(t);
    const float p = __builtin_amdgcn_exp2f((float)(t - n));
    return ldexp((double)p, (int)fmax(n, -2000.0));
}

template <bool FAST = false>
__device__ __forceinline__ void radial_parts(const EvalArgs &a, const DevShell &sh, double s,
                                             double &g, double &h) {
    g = 0.0;
    h = 0.0;
    for (int p0 = 0; p0 < sh.nprim; p0 += 4) {
        double xp[4], cc[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const bool in = p0 + k < sh.nprim;
            xp[k] = in ? a.exps[sh.poff + p0 + k] : 0.0;
            cc[k] = in ? a.ncc[sh.poff + p0 + k] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (p0 + k < sh.nprim) {
                const double q = cc[k] * (FAST ? exp_f32acc(-xp[k] * s) : exp(-xp[k] * s));
                g += q;
                h += xp[k] * q;
            }
        }
    }
}

// Candidate lists (below) for bases of at least this many shells: with fewer,
// nearly every shell reaches every tile (tetracene, 144 shells: 6% slower
// with the list, which costs registers and a pass over the shells)
constexpr int kCandMinShells = 400;
// Shells whose candidacy the staged kernel settles per pass of its loop
constexpr int kCandChunk = 1024;
// LDS beyond the image and the block positions: candidate list, bounding
// sphere, per-wavefront counts
constexpr size_t kCandBytes = kCandChunk * 2 + 4 * 8 + 4 * 4;

// LDS image of a tile of rows for basis_eval_staged_kernel:
//   [tile_points][lda/8][3][8] BF16 parts of phi, then [3][tile_points][lda] FP32 gradients
extern __shared__ __align__(16) unsigned char basis_tile[];

struct Comp {
    signed char e[21][3];
};

__host__ __device__ constexpr Comp lex_table(int L) {
    Comp t{};
    int k = 0;
    for (int a = L; a >= 0; a--)
        for (int b = L - a; b >= 0; b--) {
            t.e[k][0] = (signed char)a;
            t.e[k][1] = (signed char)b;
            t.e[k][2] = (signed char)(L - a - b);
            k++;
        }
    return t;
}

// TURBOMOLE component order, gtodefs.f90:108-123
__host__ __device__ constexpr Comp tm_table(int L) {
    Comp t{};
    if (L == 2) {
        const signed char d[6][3] = {{2,0,0},{0,2,0},{0,0,2},{1,1,0},{1,0,1},{0,1,1}};
        for (int i = 0; i < 6; i++) for (int k = 0; k < 3; k++) t.e[i][k] = d[i][k];
    } else if (L == 3) {
        const signed char f[10][3] = {{3,0,0},{0,3,0},{0,0,3},{2,1,0},{2,0,1},
                                      {1,2,0},{0,2,1},{1,0,2},{0,1,2},{1,1,1}};
        for (int i = 0; i < 10; i++) for (int k = 0; k < 3; k++) t.e[i][k] = f[i][k];
    } else if (L == 4) {
        const signed char g[15][3] = {{4,0,0},{0,4,0},{0,0,4},{3,1,0},{3,0,1},{1,3,0},
                                      {0,3,1},{1,0,3},{0,1,3},{2,2,0},{2,0,2},{0,2,2},
                                      {2,1,1},{1,2,1},{1,1,2}};
        for (int i = 0; i < 15; i++) for (int k = 0; k < 3; k++) t.e[i][k] = g[i][k];
    } else if (L == 5) {
        const signed char h[21][3] = {{5,0,0},{0,5,0},{0,0,5},{4,1,0},{4,0,1},{1,4,0},
                                      {0,4,1},{1,0,4},{0,1,4},{3,2,0},{3,0,2},{2,3,0},
                                      {0,3,2},{2,0,3},{0,2,3},{3,1,1},{1,3,1},{1,1,3},
                                      {2,2,1},{2,1,2},{1,2,2}};
        for (int i = 0; i < 21; i++) for (int k = 0; k < 3; k++) t.e[i][k] = h[i][k];
    } else {
        t = lex_table(L);
    }
    return t;
}

// Row m - L of T_l (host_basis.cpp sph_transform) combines only monomials
// x^a y^b z^c with b odd iff m < 0, a + b of the parity of |m| and c of the
// parity of l - |m| (its terms are x^(2t+|m|-2u-v) y^(2u+v) z^(l-2t-|m|) with
// v odd iff m < 0); every other entry is zero and is skipped at compile time.
__host__ __device__ constexpr bool sph_term(int L, int mi, int a, int b, int c) {
    const int m = mi - L, am = m < 0 ? -m : m, yp = m < 0 ? 1 : 0;
    return (a & 1) == ((am + yp) & 1) && (b & 1) == yp && (c & 1) == ((L - am) & 1);
}

__host__ __device__ constexpr int sph_off(int L) {
    return L == 2 ? 0 : L == 3 ? 30 : L == 4 ? 100 : L == 5 ? 235 : 0;
}

// MODE 0: FP64 rows in global memory; 3: split values and FP32 gradients
// into the staged kernel's LDS image; 4: FP64 values and gradients into it
template <int MODE>
__device__ __forceinline__ void store(const EvalArgs &a, size_t row, int f, double v,
                                      double gx, double gy, double gz) {
    const size_t i = row * (size_t)a.lda + f;
    const size_t plane = (size_t)a.gplane;
    if (MODE == 0) {
        a.phi64[i] = v;
        a.grad64[i] = gx;
        a.grad64[plane + i] = gy;
        a.grad64[2 * plane + i] = gz;
    } else if (MODE == 3) {
        // into the LDS tile: `row` is the point's row within the tile; the
        // value as its first kAParts BF16 parts (h + m + l = x exactly)
        const int tp = a.tile_points, lda = a.lda;
        float x = (float)v;
        __bf16 *d = (__bf16 *)basis_tile + row * kAParts * (size_t)lda + (f >> 3) * (8 * kAParts) +
                    (f & 7);
#pragma unroll
        for (int k = 0; k < kAParts; k++) {
            const __bf16 part = (__bf16)x;
            d[8 * k] = part;
            x -= (float)part;
        }
        float *gr = (float *)(basis_tile + (size_t)tp * lda * (2 * kAParts)) + row * (size_t)lda + f;
        gr[0] = (float)gx;
        gr[(size_t)tp * lda] = (float)gy;
        gr[2 * (size_t)tp * lda] = (float)gz;
    } else if (MODE == 4) {
        // FP64 into the LDS tile: value plane, then the three gradient planes
        const size_t pl = (size_t)a.tile_points * a.lda;
        double *d = (double *)basis_tile + row * (size_t)a.lda + f;
        d[0] = v;
        d[pl] = gx;
        d[2 * pl] = gy;
        d[3 * pl] = gz;
    }
}

template <int MODE, int L, bool TM, bool SPH>
__device__ void eval_shell(const EvalArgs &a, size_t row, int fout, const double rr[3], double g,
                           double h) {
    constexpr int NC = (L + 1) * (L + 2) / 2;
    constexpr int NS = 2 * L + 1;
    constexpr Comp T = TM ? tm_table(L) : lex_table(L);
    constexpr Comp LX = lex_table(L);  // column order of c_sph
    double px[L + 1], py[L + 1], pz[L + 1];
    px[0] = py[0] = pz[0] = 1.0;
#pragma unroll
    for (int i = 1; i <= L; i++) {
        px[i] = px[i - 1] * rr[0];
        py[i] = py[i - 1] * rr[1];
        pz[i] = pz[i - 1] * rr[2];
    }
    constexpr bool TRANSFORM = SPH && L >= 2;
    double acc[TRANSFORM ? NS : 1][4];
    if (TRANSFORM) {
#pragma unroll
        for (int m = 0; m < NS; m++) acc[m][0] = acc[m][1] = acc[m][2] = acc[m][3] = 0.0;
    }
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int ea = T.e[c][0], eb = T.e[c][1], ec = T.e[c][2];
        const double mono = px[ea] * py[eb] * pz[ec];
        const double v = mono * g;
        const double dx = (ea ? ea * px[ea ? ea - 1 : 0] * py[eb] * pz[ec] * g : 0.0) -
                          2.0 * rr[0] * mono * h;
        const double dy = (eb ? eb * px[ea] * py[eb ? eb - 1 : 0] * pz[ec] * g : 0.0) -
                          2.0 * rr[1] * mono * h;
        const double dz = (ec ? ec * px[ea] * py[eb] * pz[ec ? ec - 1 : 0] * g : 0.0) -
                          2.0 * rr[2] * mono * h;
        if (TRANSFORM) {
#pragma unroll
            for (int m = 0; m < NS; m++) {
                if (!sph_term(L, m, LX.e[c][0], LX.e[c][1], LX.e[c][2])) continue;
                const double t = c_sph[sph_off(L) + m * NC + c];
                acc[m][0] += t * v;
                acc[m][1] += t * dx;
                acc[m][2] += t * dy;
                acc[m][3] += t * dz;
            }
        } else {
            store<MODE>(a, row, fout + c, v, dx, dy, dz);
        }
    }
    if (TRANSFORM) {
#pragma unroll
        for (int m = 0; m < NS; m++)
            store<MODE>(a, row, fout + m, acc[m][0], acc[m][1], acc[m][2], acc[m][3]);
    }
}

// LMAX: the largest l the kernel instance handles (register use grows with
// it; the launcher picks the instance for the basis)
template <int MODE, bool TM, bool SPH, int LMAX = 5>
__device__ void dispatch_l(const EvalArgs &a, size_t row, const DevShell &sh, const double rr[3],
                           double g, double h) {
    switch (sh.l) {
        case 0: eval_shell<MODE, 0, TM, SPH>(a, row, sh.fout, rr, g, h); break;
        case 1: eval_shell<MODE, 1, TM, SPH>(a, row, sh.fout, rr, g, h); break;
        case 2: if (LMAX >= 2) eval_shell<MODE, (LMAX >= 2 ? 2 : 0), TM, SPH>(a, row, sh.fout, rr, g, h); break;
        case 3: if (LMAX >= 3) eval_shell<MODE, (LMAX >= 3 ? 3 : 0), TM, SPH>(a, row, sh.fout, rr, g, h); break;
        case 4: if (LMAX >= 4) eval_shell<MODE, (LMAX >= 4 ? 4 : 0), TM, SPH>(a, row, sh.fout, rr, g, h); break;
        default: if (LMAX >= 5) eval_shell<MODE, (LMAX >= 5 ? 5 : 0), TM, SPH>(a, row, sh.fout, rr, g, h); break;
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void basis_eval_kernel(EvalArgs a) {
    const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long total = (long long)a.rows * a.nshell;
    if (gid >= total) return;
    const size_t row = (size_t)(gid / a.nshell);
    const DevShell sh = a.shells[gid % a.nshell];
    bool zero = row >= (size_t)a.npts;
    double rr[3] = {0.0, 0.0, 0.0}, g = 0.0, h = 0.0;
    if (!zero) {
        const double *r = a.pts + 3 * row;
        const double *R = a.coords + 3 * sh.atom;
        rr[0] = r[0] - R[0];
        rr[1] = r[1] - R[1];
        rr[2] = r[2] - R[2];
        const double s = rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2];
        zero = sqrt(s) > sh.radius;
        if (!zero) radial_parts(a, sh, s, g, h);
    }
    if (zero) {
        for (int f = 0; f < sh.ncomp_out; f++) store<MODE>(a, row, sh.fout + f, 0.0, 0.0, 0.0, 0.0);
        return;
    }
    if (a.turbomole) {
        if (a.spherical) dispatch_l<MODE, true, true>(a, row, sh, rr, g, h);
        else dispatch_l<MODE, true, false>(a, row, sh, rr, g, h);
    } else {
        if (a.spherical) dispatch_l<MODE, false, true>(a, row, sh, rr, g, h);
        else dispatch_l<MODE, false, false>(a, row, sh, rr, g, h);
    }
}

// Split output through LDS (see the file comment).  Grid: rows / tile_points
// workgroups; tile_points divides 128 (the smallest row tile of the
// contraction kernels) and the LDS image is
// tile_points * lda * 18 bytes (then, compressed, 4 nblk shorts).
//
// Compressed output (a.toff non-null): the workgroup's rows belong to row
// tile t of a.tile_rows rows, which keeps only its live 8-function groups.
// A shell is evaluated only if its first group is live: a shell with a
// function in a group that is not lies beyond its radius for every point
// of the tile (the groups of every shell within reach of the tile are
// live), and a shell that is within reach has all its functions in live
// groups, adjacent in the list, so its functions stay contiguous in the
// compressed columns.
//
// F64: FP64 values and gradients of dense rows (the FP64 contraction's
// input), 32 bytes per function in the image; no compression.  Staging
// keeps the global stores contiguous whatever the function order: a
// wavefront's shells (ordered by l) need not own adjacent columns.
// (four wavefronts per SIMD for l <= 3: the LDS images of four workgroups
// fit a CU, and the registers are held to 128)
template <int LMAX, bool F64 = false, bool CAND = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LMAX <= 3 ? 4 : 1)))
void basis_eval_staged_kernel(EvalArgs a) {
    constexpr int kFunBytes = F64 ? 32 : 2 * kAParts + 12;  // image bytes per function
    const int tid = threadIdx.x, tp = a.tile_points;
    const size_t row0 = (size_t)blockIdx.x * tp;
    int wt = a.lda;  // functions per row of this tile's output
    short *gpos = nullptr;  // place of each 8-function group in the compressed rows
    __bf16 *phi_dst = a.phis + row0 * kAParts * (size_t)a.lda;
    float *grad_dst = a.grad32 + row0 * (size_t)a.lda;
    size_t grad_plane = (size_t)a.gplane;
    if (!F64 && a.toff) {
        const int t = (int)(row0 / a.tile_rows);
        const int kg = a.k8count[t];
        if (kg == 0) return;  // nothing of this tile is read (uniform per workgroup)
        wt = 8 * kg;
        gpos = (short *)(basis_tile + a.lds_image);
        for (int i = tid; i < 4 * a.nblk; i += 256) gpos[i] = -1;
        __syncthreads();
        for (int i = tid; i < kg; i += 256) gpos[a.k8list[(size_t)t * a.nblk * 4 + i] >> 12] = (short)i;
        const size_t base = (size_t)a.toff[t] * a.tile_rows * 8;
        const size_t rin = row0 - (size_t)t * a.tile_rows;
        phi_dst = a.phis + base * kAParts + rin * kAParts * (size_t)wt;
        grad_dst = a.grad32 + base * 3 + rin * (size_t)wt;
        grad_plane = (size_t)a.tile_rows * wt;
    }
    // Candidate shells (CAND): the shells that can reach a row of this workgroup,
    // listed in shell order, kCandChunk at a time.  A shell is dropped only
    // if its group is not live (compressed output) or if it is beyond its
    // screening radius from the rows' bounding sphere, hence from every row;
    // the per-row radius test below still decides each (row, shell) pair,
    // so the output is the same, but the lanes of a wavefront no longer
    // idle beside pairs that every row screens out.
    short *cand = (short *)(basis_tile + a.lds_image + ((size_t)a.nblk * 8 + 15) / 16 * 16);
    double *sphere = (double *)(cand + kCandChunk);  // centre x, y, z and radius
    int *wcnt = (int *)(sphere + 4);                 // kept shells per wavefront
    if (CAND) {
        // the rows' bounding sphere: centre = mean of the valid rows, radius
        // = the largest distance from it (wavefront 0, one row per lane)
        if (tid < 64) {
            const bool v = tid < tp && row0 + tid < (size_t)a.npts;
            double p[3], c[3], n = v ? 1.0 : 0.0;
#pragma unroll
            for (int d = 0; d < 3; d++) c[d] = p[d] = v ? a.pts[3 * (row0 + tid) + d] : 0.0;
#pragma unroll
            for (int off = 32; off; off >>= 1) {
#pragma unroll
                for (int d = 0; d < 3; d++) c[d] += __shfl_xor(c[d], off);
                n += __shfl_xor(n, off);
            }
            double r = 0.0;
            if (v) {
                double q = 0.0;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const double t = p[d] - c[d] / n;
                    q += t * t;
                }
                r = sqrt(q);
            }
#pragma unroll
            for (int off = 32; off; off >>= 1) r = fmax(r, __shfl_xor(r, off));
            if (tid == 0) {
                for (int d = 0; d < 3; d++) sphere[d] = n > 0.0 ? c[d] / n : 0.0;
                sphere[3] = n > 0.0 ? r : -1.0;  // negative: no valid row
            }
        }
        __syncthreads();
    }
    const int lane = tid & 63, wave = tid >> 6;
    // fills cand[] with the kept shells of [s0, s0 + kCandChunk); returns their count
    auto candidates = [&](int s0) -> int {
        int n = 0;
        for (int i0 = 0; i0 < kCandChunk && s0 + i0 < a.nshell; i0 += 256) {
            const int si = s0 + i0 + tid;
            bool keep = false;
            if (i0 + tid < kCandChunk && si < a.nshell && sphere[3] >= 0.0) {
                const DevShell sh = a.shells[si];
                keep = !gpos || gpos[sh.fout >> 3] >= 0;
                if (keep) {
                    const double *R = a.coords + 3 * sh.atom;
                    const double dx = sphere[0] - R[0], dy = sphere[1] - R[1], dz = sphere[2] - R[2];
                    const double dist = sqrt(dx * dx + dy * dy + dz * dz);
                    // margin: beyond the radius by more than rounding can explain
                    keep = dist - sphere[3] <= sh.radius * (1.0 + 1e-12) + 1e-12;
                }
            }
            const unsigned long long m = __ballot(keep);
            if (lane == 0) wcnt[wave] = __popcll(m);
            __syncthreads();
            int off = n;
            for (int w = 0; w < wave; w++) off += wcnt[w];
            if (keep) cand[off + __popcll(m & ((1ull << lane) - 1))] = (short)(i0 + tid);
            for (int w = 0; w < 4; w++) n += wcnt[w];
            __syncthreads();
        }
        return n;
    };
    // one chunk holds every shell of most bases: then the list is built once
    // for all passes.  The FP64 rows (one or two per workgroup) keep the
    // plain walk over every shell: for so few rows the list costs more than
    // the idle lanes it saves (8.2 ms against 9.7 ms on the 60-atom cage)
    constexpr bool kCand = CAND && !F64;
    const bool one_chunk = !kCand || a.nshell <= kCandChunk;
    const int ncand1 = kCand && one_chunk ? candidates(0) : 0;
    // rows per pass: as many as the LDS image holds at this width (at least
    // one: the image holds a dense row)
    const int rpp = max(1, min(tp, (int)(a.lds_image / ((size_t)wt * kFunBytes))));
    for (int r0 = 0; r0 < tp; r0 += rpp) {
        const int nr = min(rpp, tp - r0);
        EvalArgs b = a;
        b.lda = wt;         // row stride of the LDS image
        b.tile_points = nr; // rows in the image
        const size_t phi_bytes = (size_t)nr * wt * (F64 ? 8 : 2 * kAParts),
                     grad_bytes = (size_t)nr * wt * (F64 ? 24 : 12);
        {
            uint4 *z = (uint4 *)basis_tile;
            const size_t n = (phi_bytes + grad_bytes) / 16;
            for (size_t i = tid; i < n; i += 256) z[i] = make_uint4(0u, 0u, 0u, 0u);
        }
        __syncthreads();
        // one (row, shell) pair: shell-major, the nr rows of one shell are
        // neighbouring lanes, which load the same shell and primitive data
        auto pair = [&](int lp, int si) __attribute__((always_inline)) {
            const size_t row = row0 + r0 + lp;
            if (row >= (size_t)a.npts) return;  // padding rows stay zero
            DevShell sh = a.shells[si];
            if (gpos) {
                const int cg = gpos[sh.fout >> 3];
                if (cg < 0) return;  // zero over the whole tile
                sh.fout = cg * 8 + (sh.fout & 7);
            }
            const double *r = a.pts + 3 * row;
            const double *R = a.coords + 3 * sh.atom;
            double rr[3] = {r[0] - R[0], r[1] - R[1], r[2] - R[2]};
            const double s = rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2];
            if (sqrt(s) > sh.radius) return;  // screened: zeros already in place
            double g, h;
            radial_parts<!F64>(a, sh, s, g, h);
            constexpr int M = F64 ? 4 : 3;
            if (a.turbomole) {
                if (a.spherical) dispatch_l<M, true, true, LMAX>(b, lp, sh, rr, g, h);
                else dispatch_l<M, true, false, LMAX>(b, lp, sh, rr, g, h);
            } else {
                if (a.spherical) dispatch_l<M, false, true, LMAX>(b, lp, sh, rr, g, h);
                else dispatch_l<M, false, false, LMAX>(b, lp, sh, rr, g, h);
            }
        };
        if constexpr (kCand) {
            for (int s0 = 0; s0 < a.nshell; s0 += kCandChunk) {
                const int ncand = one_chunk ? ncand1 : candidates(s0);
                for (int it = tid; it < nr * ncand; it += 256) pair(it % nr, s0 + cand[it / nr]);
                if (!one_chunk) __syncthreads();  // cand[] is read out before the next chunk refills it
            }
        } else {
            for (int it = tid; it < nr * a.nshell; it += 256) pair(it % nr, it / nr);
        }
        __syncthreads();
        // this pass's rows are contiguous in both outputs
        if (F64) {
            const size_t n16 = (size_t)nr * wt / 2;  // 16-byte units per plane
            for (int d = 0; d < 4; d++) {
                const uint4 *src = (const uint4 *)(basis_tile + (size_t)d * nr * wt * 8);
                double *base = d == 0 ? a.phi64 : a.grad64 + (size_t)(d - 1) * a.gplane;
                uint4 *dst = (uint4 *)(base + (row0 + r0) * (size_t)wt);
                for (size_t i = tid; i < n16; i += 256) dst[i] = src[i];
            }
            __syncthreads();
            continue;
        }
        {
            const uint4 *src = (const uint4 *)basis_tile;
            uint4 *dst = (uint4 *)(phi_dst + (size_t)r0 * kAParts * wt);
            for (size_t i = tid; i < phi_bytes / 16; i += 256) dst[i] = src[i];
        }
        for (int d = 0; d < 3; d++) {
            const uint4 *src = (const uint4 *)(basis_tile + phi_bytes + (size_t)d * nr * wt * 4);
            uint4 *dst = (uint4 *)(grad_dst + d * grad_plane + (size_t)r0 * wt);
            for (size_t i = tid; i < (size_t)nr * wt / 4; i += 256) dst[i] = src[i];
        }
        __syncthreads();  // the image is read out before the next pass clears it
    }
}

hipError_t upload_basis_tables(const double *sph, int nsph, hipStream_t s) {
    if (nsph > 466) return hipErrorInvalidValue;
    return hipMemcpyToSymbolAsync(HIP_SYMBOL(c_sph), sph, sizeof(double) * nsph, 0,
                                  hipMemcpyHostToDevice, s);
}

hipError_t launch_basis_eval(const EvalArgs &a, int mode, hipStream_t s) {
    const long long total = (long long)a.rows * a.nshell;
    const unsigned nblk = (unsigned)((total + 255) / 256);
    if (total <= 0) return hipSuccess;
    if (mode == 0 && (size_t)a.lda * 32 <= 150 * 1024 && a.lmax >= 0 && a.lmax <= kDevMaxL) {
        // FP64 rows staged through LDS: the largest tile of 1, 2, 4 or 8 rows
        // within 72 KiB (or one row up to 150 KiB)
        EvalArgs b = a;
        const size_t row_bytes = (size_t)a.lda * 32;
        int tp = 8;
        while (tp > 1 && tp * row_bytes > 72 * 1024) tp >>= 1;
        if (a.rows % tp) return hipErrorInvalidValue;
        b.tile_points = tp;
        b.lds_image = tp * row_bytes;
        b.toff = nullptr;
        b.nblk = 0;  // no block positions: the candidate list follows the image
        const size_t lds = b.lds_image;
        auto launch = [&](auto kernel) -> hipError_t {
            hipError_t e = hipFuncSetAttribute((const void *)kernel,
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)std::max<size_t>(lds, 72 * 1024));
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(kernel, dim3((unsigned)(a.rows / tp)), dim3(256), lds, s, b);
            return hipGetLastError();
        };
        if (a.lmax <= 1) return launch(basis_eval_staged_kernel<1, true>);
        if (a.lmax == 2) return launch(basis_eval_staged_kernel<2, true>);
        if (a.lmax == 3) return launch(basis_eval_staged_kernel<3, true>);
        return launch(basis_eval_staged_kernel<5, true>);
    }
    if (mode == 0)
        hipLaunchKernelGGL(basis_eval_kernel<0>, dim3(nblk), dim3(256), 0, s, a);
    else {
        // dense: the largest tile of 1, 2, 4 or 8 rows whose LDS image fits in
        // 72 KiB (two workgroups per CU).  Compressed: 32 rows per workgroup
        // staged in passes through an image of 36 KiB or one dense row,
        // whichever is larger (four workgroups per CU at nbf <= 2048); 32
        // rows rather than 8 spread the workgroup's fixed work (block
        // positions, candidate list) over more rows: 6-8% on the cage's and
        // the tetracene scan's basis evaluation, 64 rows no better
        EvalArgs b = a;
        const size_t row_bytes = (size_t)a.lda * (2 * kAParts + 12);
        int tp = a.toff ? 32 : 8;
        size_t image;
        if (a.toff) {
            image = std::max<size_t>(36 * 1024, row_bytes);
        } else {
            while (tp > 1 && tp * row_bytes > 72 * 1024) tp >>= 1;
            image = tp * row_bytes;
        }
        if (row_bytes > 150 * 1024) return hipErrorInvalidValue;
        b.tile_points = tp;
        b.lds_image = image;
        // the image, the group positions (16-byte aligned) and, for bases of
        // kCandMinShells shells or more, the candidate list
        const bool cand = a.nshell >= kCandMinShells;
        const size_t lds = image + ((size_t)a.nblk * 8 + 15) / 16 * 16 + (cand ? kCandBytes : 0);
        // (the candidate list's bounding sphere takes one row per lane of a wavefront)
        if (a.rows % tp || tp > 64 || a.lmax < 0 || a.lmax > kDevMaxL) return hipErrorInvalidValue;
        if (a.toff && (a.tile_rows % tp || a.rows % a.tile_rows || a.nblk * 32 != a.lda))
            return hipErrorInvalidValue;
        auto launch = [&](auto kernel) -> hipError_t {
            hipError_t e = hipFuncSetAttribute((const void *)kernel,
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)std::max<size_t>(lds, 72 * 1024));
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(kernel, dim3((unsigned)(a.rows / tp)), dim3(256), lds, s, b);
            return hipGetLastError();
        };
        hipError_t e;
        if (cand) {
            if (a.lmax <= 1) e = launch(basis_eval_staged_kernel<1, false, true>);
            else if (a.lmax == 2) e = launch(basis_eval_staged_kernel<2, false, true>);
            else if (a.lmax == 3) e = launch(basis_eval_staged_kernel<3, false, true>);
            else e = launch(basis_eval_staged_kernel<5, false, true>);
        } else {
            if (a.lmax <= 1) e = launch(basis_eval_staged_kernel<1>);
            else if (a.lmax == 2) e = launch(basis_eval_staged_kernel<2>);
            else if (a.lmax == 3) e = launch(basis_eval_staged_kernel<3>);
            else e = launch(basis_eval_staged_kernel<5>);
        }
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

}  // namespace gimic
