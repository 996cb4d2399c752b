/*
 * CPU FP64 restatement of the reference GIMIC current-density tensor.
 * TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the bench
 * CPU baseline as the checker, never by the product path.
 *
 * Basis description (segmented contractions, one entry per contraction in
 * the reference's function order: atom-major, contractions as read):
 *   shell_atom[s], shell_l[s], shell_nprim[s], shell_poff[s] (offset into
 *   exps/coefs), raw contraction coefficients (normalised here exactly as
 *   basis.f90:164-191 norm_ctr does).
 * Densities: nset (1 closed shell, 2 alpha+beta) blocks of 4 n*n
 *   column-major matrices (D, DBx, DBy, DBz), used as given.
 * Output: 9 doubles per point, ct(m,b) column-major (jtensor.F90:148-237).
 */
#ifndef GIMIC_JTENSOR_ORACLE_H
#define GIMIC_JTENSOR_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* spin_case: 0 total, 1 alpha, 2 beta, 3 spindens (jtensor.F90:66-103) */
int oracle_jtensor(int natoms, const double *coords, int nshell,
                   const int *shell_atom, const int *shell_l,
                   const int *shell_nprim, const int *shell_poff,
                   const double *exps, const double *coefs, int turbomole,
                   double screen_thrs, int nbf, const double *dens, int nset,
                   int spin_case, int giao, int diamag, int paramag,
                   int64_t npts, const double *pts, double *out);

/* basis values (nbf) and gradients (3*nbf, axis-major) at one point
 * (bfeval.f90:81-122, dfdr 295-338) */
int oracle_basis(int natoms, const double *coords, int nshell,
                 const int *shell_atom, const int *shell_l,
                 const int *shell_nprim, const int *shell_poff,
                 const double *exps, const double *coefs, int turbomole,
                 double screen_thrs, const double *r, double *phi,
                 double *grad);

/* normalised coefficients (norm_ctr) and screening radii (setup_screening) */
int oracle_normalise(int nshell, const int *shell_l, const int *shell_nprim,
                     const int *shell_poff, const double *exps,
                     const double *coefs, double screen_thrs, double *ncc,
                     double *radius);

#ifdef __cplusplus
}
#endif
#endif
