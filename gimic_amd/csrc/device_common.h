// Shared device-side definitions for the CDNA4 (gfx950) current-density kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace gimic {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

// Number of per-point partial sums produced by the contraction epilogue:
//   0        rho   = sum_nu Y0 phi
//   1..3     E_c   = sum_nu Y0 phi (r - R_nu)_c   (= rho r_c - C_c)
//   4..12    A_mb  = sum_nu Yb_b G_m              (index 4 + m + 3b)
//   13..21   Q_md  = sum_nu G_m Yw_d              (index 13 + m + 3d)
// with Y0 = Phi D, Yb_b = Phi DB_b, Yw_d = Phi W_d, W_d(mu,nu) =
// D(mu,nu) (R_nu - R_mu)_d, R_nu the centre of function nu and r the point.
constexpr int kNQ = 22;

// Matrices in the contraction's right operand, per density set.
constexpr int kNMat = 7;  // D, DBx, DBy, DBz, Wx, Wy, Wz

// Point rows per workgroup tile and columns per N-tile of the contraction.
constexpr int kWgRows = 128;
constexpr int kNT = 16;
// K padding granularity (FP64 K-step 16, BF16 K-step 32)
constexpr int kKPad = 32;

struct DevShell {
    int atom;
    int l;
    int nprim;
    int poff;
    int fout;       // first output function
    int ncomp_out;  // output functions of the shell
    double radius;  // screening radius
};

// Largest l handled by the device kernels (globals.f90:30 MAX_L = 5)
constexpr int kDevMaxL = 5;

__device__ __forceinline__ uint16_t f32_to_bf16_rn(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __uint_as_float(((uint32_t)h) << 16);
}

// Sum over the 16 lanes of a DPP row (lanes 16j .. 16j+15), left in every
// lane: partners 1 and 2 apart within each quad, then the half-row mirror
// (lane i with 7 - i) joins quads 0-1 and 2-3, the row mirror (i with
// 15 - i) the two halves.  Each step adds the same pairs as a xor
// butterfly over distances 1, 2, 4, 8 (the lanes of a quad, of a half-row,
// already agree), so the sums are bit for bit those of the butterfly, with
// one VALU instruction per step and no LDS crossbar traffic.
#define GIMIC_DPP(x, ctrl) \
    __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), (ctrl), 0xF, 0xF, false))
__device__ __forceinline__ float row16_sum(float x) {
    x += GIMIC_DPP(x, 0xB1);   // quad_perm [1,0,3,2]
    x += GIMIC_DPP(x, 0x4E);   // quad_perm [2,3,0,1]
    x += GIMIC_DPP(x, 0x141);  // row_half_mirror
    x += GIMIC_DPP(x, 0x140);  // row_mirror
    return x;
}
#define GIMIC_DPP_D(x, ctrl)                                                                    \
    __hiloint2double(__builtin_amdgcn_update_dpp(0, __double2hiint(x), (ctrl), 0xF, 0xF, false), \
                     __builtin_amdgcn_update_dpp(0, __double2loint(x), (ctrl), 0xF, 0xF, false))
__device__ __forceinline__ double row16_sum(double x) {
    x += GIMIC_DPP_D(x, 0xB1);
    x += GIMIC_DPP_D(x, 0x4E);
    x += GIMIC_DPP_D(x, 0x141);
    x += GIMIC_DPP_D(x, 0x140);
    return x;
}

// launchers (defined in the .hip files)
struct EvalArgs {
    const double *pts;   // 3*npts of this chunk
    int npts;            // valid points
    int rows;            // padded rows (multiple of kWgRows)
    long long gplane;    // elements between the three gradient planes
    const DevShell *shells;
    int nshell;
    const double *exps, *ncc, *coords;
    int turbomole, spherical;
    int lda;             // padded K (multiple of kKPad)
    // FP64 outputs
    double *phi64, *grad64;  // [rows][lda], [3][rows][lda]
    // mixed outputs: FP32 gradients, and each value x as kAParts BF16 parts
    // (x = h + m + l exactly with three), layout [row][lda/8][kAParts][8]
    // (compressed to the live groups of each row tile, below)
    float *grad32;
    __bf16 *phis;
    int tile_points;     // points per workgroup of the LDS-staged kernel (set by the launcher)
    size_t lds_image;    // bytes of its LDS image (set by the launcher)
    int lmax;            // largest shell l of the basis
    // Compressed split output (staged kernel, toff non-null): a row tile of
    // tile_rows rows keeps only its k8count[t] live 8-function groups
    // (k8list[t * 4 nblk ...], global group << 12 | place k in the row); its
    // rows start at element toff[t] * tile_rows * 8 kAParts of phis
    // ([row][k][kAParts][8]) and toff[t] * tile_rows * 24 of grad32
    // ([3][tile_rows][8 k])
    const int *k8count, *k8list, *toff;
    int nblk, tile_rows;
};
hipError_t launch_basis_eval(const EvalArgs &a, int mode, hipStream_t s);
hipError_t upload_basis_tables(const double *sph, int nsph, hipStream_t s);

struct ContractArgs {
    int rows, lda, ncol_pad;
    long long gplane;    // elements between the three gradient planes
    int ntiles, nsplit, tiles_per_split;
    int nmatset;      // density sets packed side by side (7 matrices each)
    const double *pts;   // chunk points (3 per row, npts valid rows)
    int npts;
    const double *a64, *g64, *b64;  // FP64 path
    const float *g32;    // mixed path: FP32 gradients
    const __bf16 *as3;   // split basis values [rows][lda/8][kAParts][8] (mixed kernels)
    const __bf16 *bs3;   // split right operand, pair blocks per (set, K-group, N-group)
    const double *rf64;  // [3][ncol_pad] atom coordinate of each function
    const float *rf32;
    double *part;        // [nmatset][nsplit][kNQ][rows]; pre-split kernels: [nmatset][nsplit][rows][kNQ]
    // screening lists per row tile (TileArgs): the live 8-function groups of
    // the listed 32-function blocks, k8count[t] of them in k8list[t * 4 nblk
    // ...]; the pre-split kernels take four per N-tile and K-step, the FP64
    // kernel two (a zeroed 12 KiB buffer stands in for its padding groups);
    // split s of a tile takes N-tiles [s * count / nsplit, (s+1) * count / nsplit)
    int nblk;
    const int *k8count, *k8list;
    const __bf16 *bzero;
    // compressed operands (pre-split kernels; see EvalArgs): null for dense
    const int *toff;
};
// FP64 kernel: row tiles of tile_rows = 64 or 128 points (a.rows a multiple)
hipError_t launch_contract_f64(const ContractArgs &a, int tile_rows, hipStream_t s);
// Pre-split variant: both operands arrive as BF16 parts, the right operand in
// pair blocks that the kernel gathers into its LDS buffers (see
// contract_mx2.hip); nbuf 2 or 3 LDS buffers, krb 2 or 4 row blocks per
// wavefront, 4 or 8 wavefronts (16 * krb * waves rows per workgroup: 4 waves
// with krb 2, nbuf 2 or 3, or krb 4, nbuf 2; 8 waves with krb 2, nbuf 2 or
// 3).  Writes every partial sum of a.part, row-major (no clearing needed).
hipError_t launch_contract_mx2(const ContractArgs &a, int nbuf, int krb, int waves, hipStream_t s);
// Mixed arithmetic: an FP32 operand value x is the sum of BF16 parts
// x = h + m + l (three parts: exact); the basis values (left operand) keep
// kAParts of them and the right operand kBParts, and a block product keeps
// the part products of split order <= 2 (h = 0, m = 1, l = 2)
constexpr int kAParts = 3;
constexpr int kBParts = 3;
constexpr int split_products() {
    int n = 0;
    for (int i = 0; i < kAParts; i++)
        for (int j = 0; j < kBParts; j++) n += i + j <= 2;
    return n;
}
constexpr int kMx2Cols = 224;                             // 7 matrices x 32 functions
// right operand in HBM: one pair block per (density set, 8-function K-group
// gk, 8-function N-group gn) of kBParts parts x 7 matrices x 8 N-functions x
// 8 K-functions, [part][matrix][n][k]; K-groups major, each a row of
// ncol_pad / 8 + 1 pair blocks whose last is zero, and a zero row last
// (lda / 8 + 1 rows), so that the padding entries of the group lists
// (-1) address zeros
constexpr int kMx2PairElems = kBParts * kNMat * 64;      // 1344 for three parts
__host__ __device__ constexpr size_t mx2_set_elems(int lda, int ncol_pad) {
    return (size_t)(lda / 8 + 1) * (size_t)(ncol_pad / 8 + 1) * kMx2PairElems;
}
// the kernel's LDS buffer of one K-step: four K-groups (slots), each
// kBParts parts x 224 columns x 8, [slot][part][column][8], column = matrix
// * 32 + 8 * (N-group of the N-tile) + function; in whole KiB
constexpr int kMx2GroupElems = kBParts * kMx2Cols * 8;   // 5376 for three parts
constexpr int kMx2BlockElems = (4 * kMx2GroupElems * 2 + 1023) / 1024 * 512;

struct FinalizeArgs {
    const double *part;  // [nmatset][nsplit][kNQ][rows], or [nmatset][nsplit][rows][kNQ] if part_rows
    int part_rows;
    int nsplit, rows, npts, nmatset;
    const double *pts;   // chunk points in row order
    const int *perm;     // caller index (within the chunk) of each row, or null
    const long long *pidx;  // output index of each caller index (subset runs), or null
    int giao, diamag, paramag;
    double *jt;          // [nmatset][npts][9] (chunk offset applied)
    long long jt_stride; // elements between matrix sets in jt
    // flux (optional): contribution w_p * n_p . (J B_p) per point
    const double *w;     // weights or null
    const int *plane;    // plane index per point or null (single plane 0)
    const double *pnorm, *pbfield, *pcenter, *pradius;  // per plane
    double *contrib;     // [npts] or null
    double *contrib_mod; // signed-modulus contribution w_p sgn(n.jv)|jv| or null
    double *contrib_acid; // ACID contribution w_p dT^2 or null
};
hipError_t launch_finalize(const FinalizeArgs &a, hipStream_t s);
// Shielding (per nucleus) and magnetizability sums per point partition:
// out[(k * npart + p) * 5 + q], k < nnuc shielding of nucleus k, k == nnuc
// magnetizability; q = sum w s_x, s_y, s_z, positive part, negative part.
hipError_t launch_properties(const double *jt, const double *r, const double *w,
                             const long long *offs, const double *nuc, int nnuc, int npart,
                             double *out, hipStream_t s);
// p[0..n) = 0 (a kernel, so it may be captured into a graph)
hipError_t launch_fill_zero(double *p, size_t n, hipStream_t s);

// Points within `radius` of any nucleus (refine.hip): idx[0..count) = their
// indices in increasing order (count written to *count on the device);
// scratch holds near_select_scratch(npts) ints.
size_t near_select_scratch(long long npts);
hipError_t launch_near_select(const double *r, long long npts, const double *coords, int natoms,
                              double radius, long long *idx, long long *count, int *scratch,
                              hipStream_t s);
// out[3j..] = r[3 idx[j]..] for j < n
hipError_t launch_gather_points(const double *r, const long long *idx, long long n, double *out,
                                hipStream_t s);

// Plane-flux inputs on the device (planes.hip): plane index of each point
// from the plane offsets, and bond-plane product grids from their
// descriptors (per plane: n1 x n2 points, origin, axes v1 v2 v3, p3 and w3;
// 1-D nodes and weights at off1 / off2; point offsets poff[nplanes + 1])
struct PlaneGridArgs {
    int nplanes;
    long long npts;
    const int *n1;
    const long long *off1, *off2, *poff;
    const double *origin, *axes, *p3w3, *nodes1, *wts1, *nodes2, *wts2;
};
hipError_t launch_plane_index(const long long *offs, int nplanes, long long npts, int *plane,
                              hipStream_t s);
hipError_t launch_plane_grid(const PlaneGridArgs &a, double *r, double *w, int *plane,
                             hipStream_t s);

// deterministic per-plane sums of contributions: out[3*plane] = net, pos, neg
hipError_t launch_plane_reduce(const double *contrib, const long long *offs, int nplanes,
                               double *out, hipStream_t s);

// Spatial tiles (tiles.hip).  launch_sort_points orders n points along a
// Morton curve of 32^3 cells over their bounding box (stable, so the same on
// every run): spts[j] = pts[perm[j]]; bpart holds 256 * 6 doubles, scratch
// sort_scratch_ints(n) ints.
int sort_scratch_ints(int n);
hipError_t launch_sort_points(const double *pts, int n, double *spts, int *perm, double *bpart,
                              int *scratch, hipStream_t s);
struct TileArgs {
    const double *spts;  // sorted chunk points
    int npts, rows;      // valid and padded rows
    int tile_rows;       // rows per contraction workgroup
    int nblk;            // 32-function blocks (lda / 32)
    const int *blk_beg;  // [nblk + 1] ranges into blk_sh
    const int *blk_sh;   // shells with a function in each block
    const DevShell *shells;
    const double *coords;
    int *bcount, *blist;  // [rows / tile_rows], [rows / tile_rows][nblk]
    // 8-function groups of the listed blocks that screening leaves non-zero:
    // k8count[t] groups, k8list[t * 4 nblk ...] = (global group << 12) |
    // (place k of the group in this list: its place in the compressed rows), padded
    // with -1 to a multiple of 4 (nblk <= 1024)
    int *k8count, *k8list;
};
hipError_t launch_block_lists(const TileArgs &a, hipStream_t s);
// toff[t] = sum of cnt[0 .. t-1] for t = 0 .. ntiles (one workgroup)
hipError_t launch_tile_offsets(const int *cnt, int ntiles, int *toff, hipStream_t s);

hipError_t run_mfma_layout_probe(double *h_f64, float *h_bf16);
hipError_t run_mfma32_layout_probe(float *h_out);  // 2048 floats

}  // namespace gimic
