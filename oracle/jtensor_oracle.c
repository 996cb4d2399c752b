/*
 * CPU FP64 restatement of the reference GIMIC current-density tensor.
 * TEST INFRASTRUCTURE ONLY (the checker; never shipped or measured as the
 * product).  Each routine follows the reference file:line given beside it.
 * Written for clarity, one point at a time, like the reference.
 */
#include "jtensor_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXL 5

/* cartesian component exponents (gtodefs.f90:86-123): default order is
 * lexicographic (a descending, then b descending); TURBOMOLE order is the
 * table from gtodefs.f90:108-123 */
static const int TM_D[6][3] = {{2,0,0},{0,2,0},{0,0,2},{1,1,0},{1,0,1},{0,1,1}};
static const int TM_F[10][3] = {{3,0,0},{0,3,0},{0,0,3},{2,1,0},{2,0,1},
                                {1,2,0},{0,2,1},{1,0,2},{0,1,2},{1,1,1}};
static const int TM_G[15][3] = {{4,0,0},{0,4,0},{0,0,4},{3,1,0},{3,0,1},
                                {1,3,0},{0,3,1},{1,0,3},{0,1,3},{2,2,0},
                                {2,0,2},{0,2,2},{2,1,1},{1,2,1},{1,1,2}};
static const int TM_H[21][3] = {{5,0,0},{0,5,0},{0,0,5},{4,1,0},{4,0,1},
                                {1,4,0},{0,4,1},{1,0,4},{0,1,4},{3,2,0},
                                {3,0,2},{2,3,0},{0,3,2},{2,0,3},{0,2,3},
                                {3,1,1},{1,3,1},{1,1,3},{2,2,1},{2,1,2},
                                {1,2,2}};

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

static void components(int l, int tm, int comp[][3])
{
    int k = 0;
    if (tm && l >= 2) {
        const int(*t)[3] = l == 2 ? TM_D : l == 3 ? TM_F : l == 4 ? TM_G : TM_H;
        for (int i = 0; i < ncart(l); i++) {
            comp[i][0] = t[i][0]; comp[i][1] = t[i][1]; comp[i][2] = t[i][2];
        }
        return;
    }
    for (int a = l; a >= 0; a--)
        for (int b = l - a; b >= 0; b--) {
            comp[k][0] = a; comp[k][1] = b; comp[k][2] = l - a - b; k++;
        }
}

/* norm_ctr, basis.f90:164-191; setup_screening, basis.f90:90-112 */
int oracle_normalise(int nshell, const int *shell_l, const int *shell_nprim,
                     const int *shell_poff, const double *exps,
                     const double *coefs, double screen_thrs, double *ncc,
                     double *radius)
{
    const double PII = 3.141592653589793;
    for (int s = 0; s < nshell; s++) {
        int l = shell_l[s], np = shell_nprim[s], o = shell_poff[s];
        if (l < 0 || l > MAXL || np < 1) return -1;
        double j = 1.0 * (l + 1), n = 0.0;
        for (int a = 0; a < np; a++)
            for (int b = 0; b <= a; b++) {
                double e1 = exps[o + a], e2 = exps[o + b];
                double t = 2.0 * sqrt(e1 * e2) / (e1 + e2);
                t = coefs[o + a] * coefs[o + b] * pow(t, j + 0.5);
                n += t;
                if (a != b) n += t;
            }
        n = 1.0 / sqrt(n);
        for (int a = 0; a < np; a++)
            ncc[o + a] = coefs[o + a] * n * pow(4.0 * exps[o + a], 0.5 * j + 0.25) *
                         pow(0.5 / PII, 0.75);
        if (screen_thrs <= 0.0) {
            radius[s] = 1.0e10;
        } else {
            double mn = 1.0e15, x = 1.0e15, dist = 0.0;
            for (int a = 0; a < np; a++)
                if (exps[o + a] < mn) mn = exps[o + a];
            while (x > screen_thrs) {
                dist += 0.25;
                x = pow(dist, l) * exp(-mn * dist * dist);
            }
            radius[s] = dist;
        }
    }
    return 0;
}

/* cgto/dcgto with cao/cao2 (caos.f90:17-110), accumulated into bfeval/dfdr
 * order (bfeval.f90:81-122, 295-338) with filter_screened (basis.f90:118-136) */
static void eval_basis(int natoms, const double *coords, int nshell,
                       const int *shell_atom, const int *shell_l,
                       const int *shell_nprim, const int *shell_poff,
                       const double *exps, const double *ncc,
                       const double *radius, int tm, const double *r,
                       double *phi, double *grad, int nbf)
{
    int comp[21][3];
    int idx = 0;
    (void)natoms;
    for (int s = 0; s < nshell; s++) {
        int l = shell_l[s], nc = ncart(l), o = shell_poff[s];
        const double *R = coords + 3 * shell_atom[s];
        double rr[3] = {r[0] - R[0], r[1] - R[1], r[2] - R[2]};
        double rr2 = rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2];
        if (sqrt(rr2) <= radius[s]) {
            double g = 0.0, h = 0.0;
            for (int p = 0; p < shell_nprim[s]; p++) {
                double q = ncc[o + p] * exp(-exps[o + p] * rr2);
                g += q;
                h += exps[o + p] * q;
            }
            components(l, tm, comp);
            for (int c = 0; c < nc; c++) {
                double mono = pow(rr[0], comp[c][0]) * pow(rr[1], comp[c][1]) *
                              pow(rr[2], comp[c][2]);
                phi[idx + c] = mono * g;
                for (int ax = 0; ax < 3; ax++) {
                    int df[3] = {comp[c][0], comp[c][1], comp[c][2]};
                    df[ax] -= 1;
                    if (df[ax] < 0) df[ax] = 0;
                    double down = comp[c][ax] * pow(rr[0], df[0]) *
                                  pow(rr[1], df[1]) * pow(rr[2], df[2]) * g;
                    double up = 2.0 * rr[ax] * mono * h;
                    grad[ax * nbf + idx + c] = down - up;
                }
            }
        } else {
            for (int c = 0; c < nc; c++) {
                phi[idx + c] = 0.0;
                grad[idx + c] = grad[nbf + idx + c] = grad[2 * nbf + idx + c] = 0.0;
            }
        }
        idx += nc;
    }
}

static int count_bf(int nshell, const int *shell_l)
{
    int n = 0;
    for (int s = 0; s < nshell; s++) n += ncart(shell_l[s]);
    return n;
}

int oracle_basis(int natoms, const double *coords, int nshell,
                 const int *shell_atom, const int *shell_l,
                 const int *shell_nprim, const int *shell_poff,
                 const double *exps, const double *coefs, int turbomole,
                 double screen_thrs, const double *r, double *phi, double *grad)
{
    int nprim = 0;
    for (int s = 0; s < nshell; s++)
        if (shell_poff[s] + shell_nprim[s] > nprim) nprim = shell_poff[s] + shell_nprim[s];
    double *ncc = malloc(sizeof(double) * (size_t)nprim);
    double *rad = malloc(sizeof(double) * (size_t)nshell);
    int rc = oracle_normalise(nshell, shell_l, shell_nprim, shell_poff, exps, coefs,
                              screen_thrs, ncc, rad);
    if (rc == 0)
        eval_basis(natoms, coords, nshell, shell_atom, shell_l, shell_nprim, shell_poff,
                   exps, ncc, rad, turbomole, r, phi, grad, count_bf(nshell, shell_l));
    free(ncc);
    free(rad);
    return rc;
}

/* contract, jtensor.F90:148-237, for one density set */
static void contract(int nbf, const double *phi, const double *grad,
                     const double *db, const double *d2, const double *D,
                     const double *r, int giao, int diamag, int paramag,
                     double *denbf, double *pdbf, double *dendb, double ct[9])
{
    const double ZETA = 0.5;
    const double *DB[3] = {D + (size_t)nbf * nbf, D + 2 * (size_t)nbf * nbf,
                           D + 3 * (size_t)nbf * nbf};
    double dpd[3], diapam = 0.0;
    /* denbf = matmul(bfvec, aodens): denbf(nu) = sum_mu phi(mu) D(mu,nu) */
    for (int nu = 0; nu < nbf; nu++) {
        double s = 0.0;
        for (int mu = 0; mu < nbf; mu++) s += phi[mu] * D[mu + (size_t)nbf * nu];
        denbf[nu] = s;
    }
    for (int nu = 0; nu < nbf; nu++) diapam += denbf[nu] * phi[nu];
    int k = 0;
    for (int b = 0; b < 3; b++) {
        for (int nu = 0; nu < nbf; nu++) {
            double s = 0.0;
            for (int mu = 0; mu < nbf; mu++) s += phi[mu] * DB[b][mu + (size_t)nbf * nu];
            pdbf[nu] = s;
        }
        if (giao) {
            for (int nu = 0; nu < nbf; nu++) {
                double s = 0.0;
                for (int mu = 0; mu < nbf; mu++)
                    s += db[b * nbf + mu] * D[mu + (size_t)nbf * nu];
                dendb[nu] = s;
            }
        }
        dpd[b] = diapam * 0.5 * r[b];
        for (int m = 0; m < 3; m++) {
            double prsp1 = 0.0, prsp2 = 0.0, ppd = 0.0;
            if (giao) {
                for (int nu = 0; nu < nbf; nu++) {
                    prsp1 -= dendb[nu] * grad[m * nbf + nu];
                    prsp2 += denbf[nu] * d2[k * nbf + nu];
                }
            }
            for (int nu = 0; nu < nbf; nu++) ppd += pdbf[nu] * grad[m * nbf + nu];
            ct[m + 3 * b] = ZETA * ppd;
            if (giao) ct[m + 3 * b] += ZETA * (prsp1 + prsp2);
            k++;
        }
    }
    if (!paramag)
        for (int i = 0; i < 9; i++) ct[i] = 0.0;
    if (!diamag) dpd[0] = dpd[1] = dpd[2] = 0.0;
    ct[0 + 3 * 1] += dpd[2];
    ct[0 + 3 * 2] -= dpd[1];
    ct[1 + 3 * 0] -= dpd[2];
    ct[1 + 3 * 2] += dpd[0];
    ct[2 + 3 * 0] += dpd[1];
    ct[2 + 3 * 1] -= dpd[0];
}

int oracle_jtensor(int natoms, const double *coords, int nshell,
                   const int *shell_atom, const int *shell_l,
                   const int *shell_nprim, const int *shell_poff,
                   const double *exps, const double *coefs, int turbomole,
                   double screen_thrs, int nbf, const double *dens, int nset,
                   int spin_case, int giao, int diamag, int paramag,
                   int64_t npts, const double *pts, double *out)
{
    if (nbf != count_bf(nshell, shell_l)) return -2;
    if (nset != 1 && nset != 2) return -3;
    if (nset == 1 && (spin_case == 2 || spin_case == 3)) return -4;
    int nprim = 0;
    for (int s = 0; s < nshell; s++)
        if (shell_poff[s] + shell_nprim[s] > nprim) nprim = shell_poff[s] + shell_nprim[s];
    double *ncc = malloc(sizeof(double) * (size_t)nprim);
    double *rad = malloc(sizeof(double) * (size_t)nshell);
    int rc = oracle_normalise(nshell, shell_l, shell_nprim, shell_poff, exps, coefs,
                              screen_thrs, ncc, rad);
    if (rc) { free(ncc); free(rad); return rc; }
    const size_t blk = 4 * (size_t)nbf * nbf;

#pragma omp parallel
    {
        double *phi = malloc(sizeof(double) * 20 * (size_t)nbf);
        double *grad = phi + nbf, *db = grad + 3 * nbf, *d2 = db + 3 * nbf;
        double *denbf = d2 + 9 * nbf, *pdbf = denbf + nbf, *dendb = pdbf + nbf;
#pragma omp for schedule(dynamic, 8)
        for (int64_t p = 0; p < npts; p++) {
            const double *r = pts + 3 * p;
            eval_basis(natoms, coords, nshell, shell_atom, shell_l, shell_nprim,
                       shell_poff, exps, ncc, rad, turbomole, r, phi, grad, nbf);
            /* mkdbop, dfdb, d2fdrdb (bfeval.f90:168-293) */
            int mu = 0;
            for (int s = 0; s < nshell; s++) {
                const double *R = coords + 3 * shell_atom[s];
                double S[3] = {r[1] * R[2] - r[2] * R[1], r[2] * R[0] - r[0] * R[2],
                               r[0] * R[1] - r[1] * R[0]};
                for (int c = 0; c < ncart(shell_l[s]); c++, mu++) {
                    double f = phi[mu], gx = grad[mu], gy = grad[nbf + mu],
                           gz = grad[2 * nbf + mu];
                    for (int b = 0; b < 3; b++) db[b * nbf + mu] = S[b] * f;
                    d2[0 * nbf + mu] = gx * S[0];
                    d2[1 * nbf + mu] = gy * S[0] + R[2] * f;
                    d2[2 * nbf + mu] = gz * S[0] - R[1] * f;
                    d2[3 * nbf + mu] = gx * S[1] - R[2] * f;
                    d2[4 * nbf + mu] = gy * S[1];
                    d2[5 * nbf + mu] = gz * S[1] + R[0] * f;
                    d2[6 * nbf + mu] = gx * S[2] + R[1] * f;
                    d2[7 * nbf + mu] = gy * S[2] - R[0] * f;
                    d2[8 * nbf + mu] = gz * S[2];
                }
            }
            double ja[9], jb[9];
            double *o = out + 9 * p;
            switch (spin_case) {
            case 1:
                contract(nbf, phi, grad, db, d2, dens, r, giao, diamag, paramag,
                         denbf, pdbf, dendb, o);
                break;
            case 2:
                contract(nbf, phi, grad, db, d2, dens + blk, r, giao, diamag, paramag,
                         denbf, pdbf, dendb, o);
                break;
            case 3:
                contract(nbf, phi, grad, db, d2, dens, r, giao, diamag, paramag,
                         denbf, pdbf, dendb, ja);
                contract(nbf, phi, grad, db, d2, dens + blk, r, giao, diamag, paramag,
                         denbf, pdbf, dendb, jb);
                for (int i = 0; i < 9; i++) o[i] = ja[i] - jb[i];
                break;
            default:
                contract(nbf, phi, grad, db, d2, dens, r, giao, diamag, paramag,
                         denbf, pdbf, dendb, ja);
                if (nset == 2) {
                    contract(nbf, phi, grad, db, d2, dens + blk, r, giao, diamag,
                             paramag, denbf, pdbf, dendb, jb);
                    for (int i = 0; i < 9; i++) o[i] = ja[i] + jb[i];
                } else {
                    memcpy(o, ja, sizeof ja);
                }
            }
        }
        free(phi);
    }
    free(ncc);
    free(rad);
    return 0;
}
