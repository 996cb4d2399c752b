/*
 * gimic_hip.h - C ABI of the MI355X (gfx950) GIMIC current-density library.
 *
 * Part 1 is the reference's own C interface, same names, same by-pointer
 * calling convention and the same tensor layout jt[m + 3*b] = J(m,b)
 * (m = current component, b = field component).  Part 2 is a batched,
 * handle-based interface that keeps the density family resident in HBM and
 * evaluates many points per call.  Plain pointers and sizes only.
 */
#ifndef GIMIC_HIP_H
#define GIMIC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Part 1: drop-in replacements (one implicit global context per process).
 * Errors are reported on stderr and through gimic_last_error(); the library
 * never terminates the process.
 * ------------------------------------------------------------------------- */

/* src/libgimic/gimic_interface.h:10, gimic_interface.f90:26-70.
 * Reads a MOL/INTGRL basis and an XDENS density file; cartesian basis,
 * screening threshold 1e-6, GIAO/diamagnetic/paramagnetic terms on, B = 0,
 * spin case "total". */
void gimic_init(const char *molfile, const char *densfile);
/* gimic_interface.h:11, gimic_interface.f90:72-78 */
void gimic_finalize(void);
/* gimic_interface.h:12, gimic_interface.f90:80-86.  Takes effect at the next
 * gimic_init (the reference's gimic_init resets it to closed shell). */
void gimic_set_uhf(int *uhf);
/* gimic_interface.h:13, gimic_interface.f90:88-91 */
void gimic_set_magnet(const double *b);
/* gimic_interface.h:14, gimic_interface.f90:93-112: "alpha", "beta",
 * "total", "spindens" */
void gimic_set_spin(const char *s);
/* gimic_interface.h:15, gimic_interface.f90:114-117.  As in the reference the
 * threshold is applied when the basis is set up, i.e. at the next gimic_init. */
void gimic_set_screening(const double *thrs);
/* gimic_interface.h:16, gimic_interface.f90:119-139 */
void gimic_calc_jtensor(const double *r, double *jt);
/* gimic_interface.h:17, gimic_interface.f90:141-151: jv = J B */
void gimic_calc_jvector(const double *r, double *jv);
/* gimic_interface.h:18, gimic_interface.f90:153-163 (the reference stops with
 * "NOT IMPLEMENTED YET"); here |J B|. */
void gimic_calc_modj(const double *r, double *mj);
/* src/libgimic/gausspoints.h:10, gausspoints.f90:12-29: composite
 * Gauss-Legendre nodes/weights on [a,b], npts points in blocks of order */
void gimic_get_gauss_points(double *a, double *b, int *npts, int *order, double *pts,
                            double *wgts);
/* gausspoints.h:11, gausspoints.c:4-7 */
void mkgausspoints(double *a, double *b, int *npts, int *order, double *pts, double *wgts);

/* ---------------------------------------------------------------------------
 * Part 2: batched handle API (no counterpart in the reference's C ABI; it
 * replaces the per-point ctensor calls of the reference's loops
 * src/fgimic/jfield.f90:114-129 and src/fgimic/integral.f90:107-165).
 * All functions return 0 on success or a negative GIMIC_E* code.
 * ------------------------------------------------------------------------- */

typedef struct gimic_ctx gimic_ctx;

enum {
    GIMIC_OK = 0,
    GIMIC_EINVAL = -1,  /* bad argument */
    GIMIC_EIO = -2,     /* file missing or malformed */
    GIMIC_EDEVICE = -3, /* HIP runtime error or no device */
    GIMIC_ENOMEM = -4,  /* device allocation failed */
    GIMIC_ESTATE = -5   /* e.g. no density set, beta requested for closed shell */
};

enum { GIMIC_SPIN_TOTAL = 0, GIMIC_SPIN_ALPHA = 1, GIMIC_SPIN_BETA = 2, GIMIC_SPIN_SPINDENS = 3 };

/* contraction precision: FP64 MFMA, or mixed: FP32 operands each split into
 * three BF16 parts with the six products of split order <= 2 on BF16 MFMA,
 * FP32 accumulation.  Basis-function exponentials, partial sums across
 * tiles, the tensor assembly and the flux sums are FP64 in both. */
enum { GIMIC_PREC_FP64 = 0, GIMIC_PREC_MIXED = 1 };

/* message of the last failing call on this thread ("" if none) */
const char *gimic_last_error(void);

/* Basis from a MOL/INTGRL file (src/libgimic/intgrl.f90:20-264), density
 * from an XDENS file (dens.f90:56-135; UHF perturbed densities halved and
 * TURBOMOLE order permuted as the reference reader does).  densfile may be
 * NULL to set densities later.  screen_thrs <= 0 disables screening. */
int gimic_ctx_create_files(gimic_ctx **ctx, const char *molfile, const char *densfile, int uhf,
                           double screen_thrs);

/* Basis from arrays.  Shells (segmented contractions) in function order:
 * atom-major, atom indices 0-based and non-decreasing; shell s has angular
 * momentum shell_l[s] <= 5 and shell_nprim[s] primitives at
 * exps/coefs[shell_poff[s] ...] (raw coefficients, normalised here as the
 * reference's norm_ctr does).  coords: 3*natoms bohr.  turbomole selects the
 * TURBOMOLE cartesian component order; spherical selects real solid
 * harmonics (2l+1 functions per shell, see gimic_sph_transform). */
int gimic_ctx_create(gimic_ctx **ctx, int natoms, const double *coords, int nshell,
                     const int *shell_atom, const int *shell_l, const int *shell_nprim,
                     const int *shell_poff, const double *exps, const double *coefs,
                     int turbomole, int spherical, double screen_thrs);

/* nset = 1 (closed shell) or 2 (alpha, beta): nset*4 blocks of nbf*nbf
 * column-major doubles D, DBx, DBy, DBz, used exactly as given. */
int gimic_ctx_set_density(gimic_ctx *ctx, int nset, const double *dens);

/* GIAO, diamagnetic and paramagnetic terms (jtensor.F90:180-228) */
int gimic_ctx_set_flags(gimic_ctx *ctx, int giao, int diamag, int paramag);

int gimic_ctx_info(const gimic_ctx *ctx, int *nbf, int *nbf_cart, int *natoms, int *nshell,
                   int *nset);

/* atom coordinates (3*natoms, bohr) and nuclear charges as read
 * (intgrl.f90:91-115); either pointer may be NULL */
int gimic_ctx_atoms(const gimic_ctx *ctx, double *coords, double *charges);

/* atom symbols as the MOL file's atom lines give them (the first two
 * characters, intgrl.f90:111-112; "X " for a basis built from arrays):
 * 2 characters per atom, not NUL-terminated, into sym[2*natoms] */
int gimic_ctx_atom_symbols(const gimic_ctx *ctx, char *sym);

/* Per-kernel timing (no counterpart in the reference): when on, every
 * launch of the basis-evaluation, contraction and finalize kernels is
 * bracketed by HIP events on its stream; gimic_ctx_get_timing waits for
 * them and returns the accumulated milliseconds and launch counts (3 each,
 * in that order).  Turning timing on or off resets the totals. */
int gimic_ctx_set_timing(gimic_ctx *ctx, int on);
int gimic_ctx_get_timing(gimic_ctx *ctx, double *ms, long long *launches);

/* Mixed-precision kernel selection (no counterpart in the reference).
 * Kernels (both operands pre-split into BF16 parts, 16x16 MFMA blocks):
 * 4 wavefronts with 128 rows per workgroup, double (3) or triple (4)
 * buffered, 4 wavefronts with 256 rows, double buffered (5), 8 wavefronts
 * with 256 rows, double (8) or triple (9) buffered.  All of them, and the
 * FP64 kernel, visit only the 32-function blocks that screening leaves
 * non-zero over their row tile.  The first mixed evaluation of a context
 * times basis evaluation + contraction on its middle chunk for kernels 3,
 * 4 and 5 (N-split count 4) and 8 and 9 (N-split counts 2, 4 and 8), each
 * twice, and keeps the fastest.  Returns the choice (variant -1 before the
 * first mixed call), its split count (0: chosen by workgroup count) and the
 * best time in ms of each variant (ms[9], index variant - 1; 0 for variants
 * not timed). */
int gimic_ctx_tuning(const gimic_ctx *ctx, int *variant, int *split, double *ms);
/* Every candidate the tuner timed: codes[k] = variant * 100 + split count,
 * ms[k] its best time.  *n: capacity in, number of candidates out. */
int gimic_ctx_tuning_log(const gimic_ctx *ctx, int *n, int *codes, double *ms);
/* Fix the mixed-precision kernel (3, 4, 5, 8 or 9, with its N-split count,
 * 0 = by workgroup count) or return to automatic selection (variant 0: the
 * next mixed evaluation tunes again). */
int gimic_ctx_set_mixed_kernel(gimic_ctx *ctx, int variant, int split);

/* Mixed precision near nuclei (no counterpart in the reference): with
 * GIMIC_PREC_MIXED, the points within `radius` bohr of any nucleus are
 * evaluated again in FP64 after the mixed pass, and their tensors and flux
 * contributions replace the mixed ones.  There the tight basis functions
 * make J large and its terms cancel, and FP32 products lose most of their
 * relative accuracy; a plane passing a nucleus at 0.1 bohr otherwise carries
 * most of its flux error in a few points.  Default 0.4 bohr; 0 turns the
 * pass off.  The selection synchronises the stream once per call.
 * gimic_ctx_fp64_refined reports the radius and how many points the last
 * mixed call evaluated in FP64 (either pointer may be NULL). */
int gimic_ctx_set_fp64_radius(gimic_ctx *ctx, double radius);
int gimic_ctx_fp64_refined(const gimic_ctx *ctx, double *radius, int64_t *npts);

/* Device memory for the basis values and gradients of one chunk of points
 * (no counterpart in the reference, which evaluates point by point).  A call
 * evaluates its points in chunks of equal size whose basis buffers fit in
 * `bytes` (mixed: 18 bytes per point and padded function; FP64: 32), and at
 * least 256 points.  0 restores the default, 8 GiB; -1 selects the large
 * setting, a third of the device's total memory and at most 96 GiB (one
 * launch for a 128^3 grid with 1800 functions).  Each launch ends in a tail
 * that leaves the GPU partly idle, so fewer, larger chunks are faster.
 * Buffers a context has allocated are kept until it is destroyed: set this
 * before the first evaluation to bound its memory.  When an allocation
 * fails, an eager call first frees the context's basis buffers of the other
 * precision, then halves its chunk.  Results of different chunk sizes agree
 * to the rounding of the precision (the tiles' block lists depend on the
 * chunk's points).  gimic_ctx_chunk_bytes reports the setting in force. */
int gimic_ctx_set_chunk_bytes(gimic_ctx *ctx, int64_t bytes);
int gimic_ctx_chunk_bytes(const gimic_ctx *ctx, int64_t *bytes);

/* Screening block-list statistics (no counterpart in the reference).  While
 * enabled (on != 0; enabling clears the totals), every eager evaluation
 * that runs a block-list kernel copies its row tiles' block counts to the
 * host (a synchronising copy: use it outside timed regions) and adds them
 * up.  out[7]: row tiles, sum of listed blocks, sum of squared listed
 * blocks, rows per tile, blocks per dense row (lda / 32), the sum over
 * tiles of listed blocks x live 8-function groups / 4, and the sum over
 * tiles of squared live groups.  A tile of R rows with k live groups
 * costs 2 * R * 7 * (8 k)^2 flops per density set (the FP64 and pre-split
 * mixed kernels walk the live groups along both N and K; with n listed
 * blocks, N-tiles of whole blocks would do 2 * R * 7 * 32 n * 8 k, and a
 * kernel that walked whole blocks both ways 2 * R * 7 * (32 n)^2). */
int gimic_ctx_set_tile_stats(gimic_ctx *ctx, int on);
int gimic_ctx_tile_stats(const gimic_ctx *ctx, double *out);

/* Graph mode for gimic_flux_planes (no counterpart in the reference): when
 * enabled, the whole plane batch (basis evaluation, contraction and tensor
 * assembly for every chunk, and the per-plane reductions) is captured once
 * into a hipGraph and launched as one graph; later calls with the same
 * sizes, flags and work buffers relaunch it without capturing again.
 * graph_stats returns the number of graph builds and launches. */
int gimic_ctx_set_graph(gimic_ctx *ctx, int enable);
int gimic_ctx_graph_stats(const gimic_ctx *ctx, long long *builds, long long *launches);

int gimic_ctx_destroy(gimic_ctx *ctx);

/* J tensors at npts points: r[3*i..] (bohr) -> jt[9*i + m + 3*b].
 * Host pointers; the call returns when the results are in jt.  More than
 * 1,376,256 points are evaluated in sub-batches whose host copies overlap
 * the kernels of their neighbours (a helper thread and a second stream of
 * the context); each sub-batch is evaluated as a call of its own, so mixed
 * results agree with a one-call device evaluation to the mixed tolerance
 * and FP64 results to summation-order rounding. */
int gimic_jtensor_batch(gimic_ctx *ctx, int64_t npts, const double *r, double *jt, int spin,
                        int precision);

/* Same with device pointers (HBM) on the given HIP stream (NULL: the
 * context's stream); asynchronous with respect to the host. */
int gimic_jtensor_batch_device(gimic_ctx *ctx, int64_t npts, const double *d_r, double *d_jt,
                               int spin, int precision, void *stream);

/* Allocation-free variant of gimic_jtensor_batch_device (no counterpart in
 * the reference; for callers that cannot allocate device memory: code
 * under HIP graph capture, tight serving loops, hosts that forbid it).
 * Same arguments, results to the rounding of the precision.  The call
 * allocates nothing: its work buffers live in a 256 MiB device array that
 * the library holds on every GPU it runs on until the process ends, and it
 * evaluates the points in chunks that fit there (any npts; a basis too wide
 * for one 256-row chunk is rejected with GIMIC_ENOMEM).  The context's
 * right operands for this spin case and precision must be resident: one
 * gimic_jtensor_batch_device call with them (one point suffices) builds
 * them, otherwise GIMIC_ESTATE.  An untuned context uses mixed kernel 9
 * with 4 N-ranges and does not tune.  Every call on one GPU shares that
 * GPU's array: calls that overlap in time on one GPU (any context, any
 * stream) race and must be serialised by the caller.  With mixed precision
 * the FP64 pass near nuclei synchronises the stream once per chunk, as in
 * the allocating call. */
int gimic_jtensor_batch_device_noalloc(gimic_ctx *ctx, int64_t npts, const double *d_r,
                                       double *d_jt, int spin, int precision, void *stream);

/* Several spin cases from one pass over the basis values: jt holds nspin
 * blocks of 9*npts doubles in the order of spins[].  Device pointers. */
int gimic_jtensor_multi_device(gimic_ctx *ctx, int64_t npts, const double *d_r, double *d_jt,
                               int nspin, const int *spins, int precision, void *stream);

/* Plane flux (src/fgimic/integral.f90:50-186) for nplanes planes in one call.
 * Points of plane k are r[3*i..] for plane_offs[k] <= i < plane_offs[k+1]
 * with product quadrature weights w[i]; per plane: unit normal, B field,
 * centre and radius bound (points farther than radius from the centre
 * contribute zero; pass 1e10 for none).  out[3*k..] = net, positive,
 * negative flux in au.  out_mod (optional, may be NULL) receives the same
 * three sums of the signed modulus of J B (integral.f90:190-324) and
 * out_acid (optional) one sum per plane of w dT^2, the ACID integrand
 * (acid.f90:11-40; the reference prints its square root, integral.f90:414-511),
 * all from the same tensors.  Host pointers; jt (optional, may be NULL)
 * receives the tensors of all points (9 per point). */
int gimic_flux_planes(gimic_ctx *ctx, int nplanes, const int64_t *plane_offs, const double *r,
                      const double *w, const double *normals, const double *bfields,
                      const double *centers, const double *radii, int spin, int precision,
                      double *out, double *out_mod, double *out_acid, double *jt);

/* Plane flux over product-quadrature planes generated on the device (the
 * bond and base grids of grid.f90, integral.f90:50-186 with the points of
 * gridpoint, grid.f90:478-495; no separate entry point in the reference).
 * Plane k has n1[k] x n2[k] points, i fastest:
 *   r = ((o + p1_i v1) + p2_j v2) + p3 v3,   w = (w3 w2_j) w1_i
 * with o = origins[3k..], v1, v2, v3 = axes[9k..], axes[9k+3..], axes[9k+6..],
 * (p3, w3) = p3w3[2k], p3w3[2k+1]; the 1-D rules of all planes are
 * concatenated in plane order (nodes1/wts1: sum n1 values, nodes2/wts2: sum
 * n2).  Each operation is rounded as written (no fused multiply-add), so the
 * points and the sums equal those of gimic_flux_planes on host-built arrays
 * bit for bit.  normals ... out_acid as for gimic_flux_planes.  Only
 * O(sum n1 + n2) numbers cross to the device; in graph mode the generator
 * runs inside the captured graph. */
int gimic_flux_plane_grids(gimic_ctx *ctx, int nplanes, const int *n1, const int *n2,
                           const double *origins, const double *axes, const double *p3w3,
                           const double *nodes1, const double *wts1, const double *nodes2,
                           const double *wts2, const double *normals, const double *bfields,
                           const double *centers, const double *radii, int spin, int precision,
                           double *out, double *out_mod, double *out_acid);

/* Magnetic properties on an external quadrature grid (jfield.f90:584-929,
 * Essential.prop with numgrid files).  J at npts points r[3*i..] with
 * weights w[i]; the points fall into npart consecutive partitions (the
 * per-atom blocks of nelpts.info), partition p = [part_offs[p],
 * part_offs[p+1]).  For each nucleus k (nuclei[3*k..], coord.au) and
 * partition p, shield[(k*npart + p)*5 + q] holds, with d = r - R_k,
 * f = -1e6 / (|d|^3 c^2), c = 137.0359998, and j_b = -J e_b:
 *   q = 0, 1, 2   sum w f (d x j_b)_b for b = x, y, z (shielding, ppm)
 *   q = 3, 4      positive and negative parts of sum w f sum_b (d x j_b)_b
 * and magnet[p*5 + q] the same sums of 1/2 (r x j_b)_b (magnetizability,
 * au).  The reference prints totals over all partitions, the isotropic
 * value (sum of the three) / 3, and per-atom contributions; the library
 * sums each partition with a fixed-order tree.  Host pointers; jt
 * (optional) receives the tensors. */
int gimic_ctx_properties(gimic_ctx *ctx, int64_t npts, const double *r, const double *w,
                         int npart, const int64_t *part_offs, int nnuc, const double *nuclei,
                         int spin, int precision, double *shield, double *magnet, double *jt);

/* Real solid-harmonic transform for angular momentum l <= 5:
 * T[(m+l)*ncart + c], m = -l..l, cartesian components in the reference's
 * non-TURBOMOLE order (gtodefs.f90:86-106); for l <= 1 the identity (s, and
 * p as x, y, z).  Rows have unit norm under the reference normalisation. */
int gimic_sph_transform(int l, double *T);

/* Composite quadrature on [a,b] as the reference's setup_gauss_data
 * (gaussint.f90:267-319): kind 0 Gauss-Legendre, 1 Gauss-Lobatto. */
int gimic_quadrature(double a, double b, int order, int npts, int kind, double *pts,
                     double *wgts);

/* ---------------------------------------------------------------------------
 * Learned surrogate of the J field (no counterpart in the reference): a
 * per-molecule coordinate network r -> 64 Fourier frequencies (sin, cos) + r
 * -> 4 x 128 SiLU -> 9 tensor elements, trained by these entry points on
 * tensors from the exact kernel (BF16 MFMA, FP32 master weights, Adam).
 * Loss: mean over points of |y - t|^2 / (|t|^2 + eps^2), t = J / out_scale.
 * Device pointers for points and tensors; the exact kernel stays the
 * fallback and the audit of its error (see gimic_amd/surrogate.py).
 * ------------------------------------------------------------------------- */
typedef struct gimic_mlp gimic_mlp;

/* sigma: std-dev of the Fourier frequencies (cycles per unit of the scaled
 * coordinate u = (r - centre)/scale); center_scale = {cx, cy, cz, scale};
 * out_scale: 9 output scales; seed: initialisation. */
int gimic_mlp_create(gimic_mlp **net, double sigma, const double *center_scale,
                     const double *out_scale, uint64_t seed);
int gimic_mlp_destroy(gimic_mlp *net);
/* run the network's kernels on `stream` (e.g. the caller's current stream),
 * NULL: on the network's own stream; waits for the previous stream first */
int gimic_mlp_set_stream(gimic_mlp *net, void *stream);
/* parameter count and the 3 x 64 frequency matrix (either may be NULL) */
int gimic_mlp_info(const gimic_mlp *net, long long *nparams, float *freqs);
/* copy the FP32 master parameters out (set = 0) or in (set = 1); layout per
 * layer l: W_l[n_out][k_in] then b_l[n_out]; k_in 160 then 128, n_out 128,
 * the output layer 16 rows (9 used) */
int gimic_mlp_params(gimic_mlp *net, float *params, int set);
/* loss and (optionally, host pointer) its gradient for one batch, no update */
int gimic_mlp_loss_grad(gimic_mlp *net, long long npts, const double *d_r,
                        const double *d_target, double eps, double *loss, float *grad);
/* one Adam step on a batch; loss (host, may be NULL) before the update */
int gimic_mlp_train_step(gimic_mlp *net, long long npts, const double *d_r,
                         const double *d_target, double eps, double lr, double *loss);
/* fused inference: d_J[9*i + m + 3*b] for npts points on `stream` (NULL: the
 * network's stream) */
int gimic_mlp_predict(gimic_mlp *net, long long npts, const double *d_r, double *d_J,
                      void *stream);
/* the same network evaluated with the training kernels (for checks) */
int gimic_mlp_predict_unfused(gimic_mlp *net, long long npts, const double *d_r, double *d_J);

/* Diagnostic (no counterpart in the reference): runs the two MFMA builtins
 * the kernels use on one wavefront with lane-numbered operands and returns
 * 512 values each (rows then columns; see gimic_amd/csrc/selftest.hip), so a
 * test can confirm the lane layout the kernels assume. */
int gimic_selftest_mfma_layout(double *f64, float *bf16);
/* The same for v_mfma_f32_32x32x16_bf16: 2048 floats, 64 lanes x 16 slots
 * with A = lane+1 (rows), then with B = lane+1 (columns). */
int gimic_selftest_mfma32_layout(float *out);

#ifdef __cplusplus
}
#endif

#endif /* GIMIC_HIP_H */
