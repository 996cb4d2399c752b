// Gauss-Legendre / Gauss-Lobatto composite quadrature (host side).
#pragma once

namespace gimic {

// Composite rule on [0, b-a] as setup_gauss_data (gaussint.f90:267-319):
// npts/order blocks of `order` nodes each; Lobatto blocks repeat their
// shared end points.  npts == 1 gives the single node 0 with weight 1.
// Throws std::runtime_error if a Newton iteration fails to converge.
void setup_gauss_data(double a, double b, int order, int npts, bool lobatto, double *pts,
                      double *wgts);

// Nodes/weights of one block on [a,b] (gaussl, gaussint.f90:21-59;
// lobattomy, gaussint.f90:61-112).
void gauss_legendre(double a, double b, int n, double *pts, double *wgts);
void gauss_lobatto(double a, double b, int n, double *pts, double *wgts);

}  // namespace gimic
