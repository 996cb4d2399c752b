// Gauss-Legendre / Gauss-Lobatto nodes by Newton iteration on Legendre
// polynomials from the three-term recurrence, as gaussint.f90 does (same
// initial guesses, tolerance 3e-12 and at most 10 Newton steps), so the
// nodes and weights agree with the reference to rounding.
#include "quadrature.h"

#include <cmath>
#include <stdexcept>
#include <vector>

namespace gimic {

namespace {

constexpr double kPi = 3.141592653589793;  // globals.f90 PII
constexpr double kEps = 3.0e-12;           // gaussint.f90:16
constexpr int kNewtonMax = 10;             // gaussint.f90:17

// P_n(x), P_n'(x), P_n''(x) by the recurrence i P_i = (2i-1) x P_{i-1} - (i-1) P_{i-2}
void legendre(double x, int n, double &y, double &dy, double &d2y) {
    y = 1.0;
    dy = 0.0;
    d2y = 0.0;
    if (n == 0) return;
    y = x;
    dy = 1.0;
    if (n == 1) return;
    double yp = 1.0, dyp = 0.0, d2yp = 0.0;
    for (int i = 2; i <= n; i++) {
        const double c1 = (double)i, c2 = c1 * 2.0 - 1.0, c4 = c1 - 1.0;
        const double ym = y;
        y = (c2 * x * y - c4 * yp) / c1;
        yp = ym;
        const double dym = dy;
        dy = (c2 * x * dy - c4 * dyp + c2 * yp) / c1;
        dyp = dym;
        const double d2ym = d2y;
        d2y = (c2 * x * d2y - c4 * d2yp + c2 * 2.0 * dyp) / c1;
        d2yp = d2ym;
    }
}

}  // namespace

void gauss_legendre(double a, double b, int n, double *pts, double *wgts) {
    const int m = (n + 1) / 2;
    const double xm = (b + a) * 0.5, xl = (b - a) * 0.5;
    for (int i = 1; i <= m; i++) {
        double z = std::cos(kPi * (i - 0.25) / (n + 0.5));
        double lp = 0.0, dlp = 0.0, d2 = 0.0;
        int it = 1;
        for (; it <= kNewtonMax; it++) {
            legendre(z, n, lp, dlp, d2);
            const double z1 = z;
            z = z1 - lp / dlp;
            if (std::fabs(z - z1) <= kEps) break;
        }
        if (it >= kNewtonMax) throw std::runtime_error("integration did not converge");
        pts[i - 1] = xm - xl * z;
        pts[n - i] = xm + xl * z;
        wgts[i - 1] = 2.0 * xl / ((1.0 - z * z) * dlp * dlp);
        wgts[n - i] = wgts[i - 1];
    }
}

void gauss_lobatto(double a, double b, int n, double *pts, double *wgts) {
    const double xm = (b + a) * 0.5, xl = (b - a) * 0.5;
    pts[0] = xm - xl;
    pts[n - 1] = xm + xl;
    wgts[0] = 2.0 * xl / (double)(n * n - n);
    wgts[n - 1] = wgts[0];
    for (int i = 2; i <= n - 1; i++) {
        double z = std::cos(kPi * (i - 0.25) / (n + 0.5));
        double lp = 0.0, dlp = 0.0, d2lp = 0.0;
        int it = 1;
        for (; it <= kNewtonMax; it++) {
            legendre(z, n - 1, lp, dlp, d2lp);
            const double z1 = z;
            z = z1 - dlp / d2lp;
            double damp = 0.5;
            while (std::fabs(z) > 1.0) {  // damped step back into [-1,1]
                z = z1 - dlp / d2lp * damp;
                damp = damp * damp;
            }
            if (std::fabs(z - z1) <= kEps) break;
        }
        if (it >= kNewtonMax) throw std::runtime_error("lobatto Newton step did not converge");
        pts[i - 1] = xm - xl * z;
        wgts[i - 1] = 2.0 * xl / ((double)(n * n - n) * lp * lp);
    }
}

void setup_gauss_data(double a, double b, int order, int npts, bool lobatto, double *pts,
                      double *wgts) {
    if (npts == 1) {
        pts[0] = 0.0;
        wgts[0] = 1.0;
        return;
    }
    if (npts % order != 0) throw std::runtime_error("npts is not dividable by ngp");
    const int nblock = npts / order;
    const double step = (b - a) / (double)nblock;
    const double xl = step * 0.5;
    std::vector<double> tp(order), tw(order);
    if (lobatto) gauss_lobatto(-1.0, 1.0, order, tp.data(), tw.data());
    else gauss_legendre(-1.0, 1.0, order, tp.data(), tw.data());
    int k = 0;
    for (int i = 0; i < nblock; i++)
        for (int j = 0; j < order; j++, k++) {
            pts[k] = (tp[j] + 1.0) * xl + (double)i * step;
            wgts[k] = tw[j] * xl;
        }
}

}  // namespace gimic
