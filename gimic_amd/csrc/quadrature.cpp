// Host-side Gauss-Legendre and Gauss-Lobatto rules (gaussint.f90:21-112,
// composite rule gaussint.f90:267-319).
//
// The nodes are roots of P_n (Legendre) or of P'_{n-1} (Lobatto interior
// nodes), found by Newton's method from the Chebyshev-like start
// cos(pi (i - 1/4) / (n + 1/2)).  The flux fixtures compare quadrature sums
// to the reference's sixth decimal and the device plane generator equals
// the host grids bit for bit, so the start values, the step tolerance, the
// step limit and the order of every product and sum follow the reference:
// then the nodes agree with it to the last bit, not just to rounding.
#include "quadrature.h"

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace gimic {

namespace {

constexpr double kPi = 3.141592653589793;  // globals.f90 PII
constexpr double kRootTol = 3.0e-12;       // Newton step that ends the search
constexpr int kMaxSteps = 10;              // a root needing the 10th step is a failure

struct LegendreValue {
    double p, dp, ddp;  // P_n(x), P_n'(x), P_n''(x)
};

// Bonnet's recurrence k P_k = (2k - 1) x P_{k-1} - (k - 1) P_{k-2} and its
// first two derivatives, differentiated term by term; each line is
// evaluated left to right as written.
LegendreValue legendre_at(int degree, double x) {
    LegendreValue cur{1.0, 0.0, 0.0};
    if (degree == 0) return cur;
    LegendreValue prev = cur;
    cur = {x, 1.0, 0.0};
    for (int k = 2; k <= degree; k++) {
        const double kk = (double)k, two_k_1 = kk * 2.0 - 1.0, k_1 = kk - 1.0;
        LegendreValue next;
        next.p = (two_k_1 * x * cur.p - k_1 * prev.p) / kk;
        next.dp = (two_k_1 * x * cur.dp - k_1 * prev.dp + two_k_1 * cur.p) / kk;
        next.ddp = (two_k_1 * x * cur.ddp - k_1 * prev.ddp + two_k_1 * 2.0 * cur.dp) / kk;
        prev = cur;
        cur = next;
    }
    return cur;
}

double start_value(int i, int n) { return std::cos(kPi * (i - 0.25) / (n + 0.5)); }

// Newton from x0: `step(x, &v)` returns the new iterate and the Legendre
// values it used; the search ends when an iterate moves by at most kRootTol.
template <class Step>
double newton_root(double x0, Step step, LegendreValue &at_root, const char *rule) {
    double x = x0;
    int n = 1;
    for (; n <= kMaxSteps; n++) {
        const double before = x;
        x = step(before, at_root);
        if (std::fabs(x - before) <= kRootTol) break;
    }
    if (n >= kMaxSteps)
        throw std::runtime_error(std::string(rule) + ": no node within " +
                                 std::to_string(kMaxSteps) + " Newton steps");
    return x;
}

}  // namespace

void gauss_legendre(double a, double b, int n, double *pts, double *wgts) {
    const double mid = (b + a) * 0.5, half = (b - a) * 0.5;
    for (int i = 1; i <= (n + 1) / 2; i++) {
        LegendreValue v{};
        const double z = newton_root(
            start_value(i, n),
            [n](double x, LegendreValue &lv) {
                lv = legendre_at(n, x);
                return x - lv.p / lv.dp;
            },
            v, "Gauss-Legendre rule");
        pts[i - 1] = mid - half * z;
        pts[n - i] = mid + half * z;
        wgts[i - 1] = 2.0 * half / ((1.0 - z * z) * (v.dp * v.dp));
        wgts[n - i] = wgts[i - 1];
    }
}

void gauss_lobatto(double a, double b, int n, double *pts, double *wgts) {
    const double mid = (b + a) * 0.5, half = (b - a) * 0.5;
    const double nn = (double)(n * n - n);
    pts[0] = mid - half;
    pts[n - 1] = mid + half;
    wgts[0] = 2.0 * half / nn;
    wgts[n - 1] = wgts[0];
    for (int i = 2; i <= n - 1; i++) {
        LegendreValue v{};
        const double z = newton_root(
            start_value(i, n),
            [n](double x, LegendreValue &lv) {
                lv = legendre_at(n - 1, x);
                // a step that leaves [-1, 1] is shortened by factors 1/2,
                // 1/4, 1/16, ... of itself until it does not
                double next = x - lv.dp / lv.ddp, shrink = 0.5;
                while (std::fabs(next) > 1.0) {
                    next = x - lv.dp / lv.ddp * shrink;
                    shrink = shrink * shrink;
                }
                return next;
            },
            v, "Gauss-Lobatto rule");
        pts[i - 1] = mid - half * z;
        wgts[i - 1] = 2.0 * half / (nn * (v.p * v.p));
    }
}

void setup_gauss_data(double a, double b, int order, int npts, bool lobatto, double *pts,
                      double *wgts) {
    if (npts == 1) {
        pts[0] = 0.0;
        wgts[0] = 1.0;
        return;
    }
    if (npts % order != 0)
        throw std::runtime_error("quadrature: " + std::to_string(npts) +
                                 " points are not a whole number of blocks of " +
                                 std::to_string(order));
    const int blocks = npts / order;
    const double width = (b - a) / (double)blocks, half = width * 0.5;
    std::vector<double> node(order), weight(order);
    if (lobatto) gauss_lobatto(-1.0, 1.0, order, node.data(), weight.data());
    else gauss_legendre(-1.0, 1.0, order, node.data(), weight.data());
    // block i covers [i width, (i + 1) width]; Lobatto blocks repeat their
    // shared end points
    for (int i = 0, k = 0; i < blocks; i++)
        for (int j = 0; j < order; j++, k++) {
            pts[k] = (node[j] + 1.0) * half + (double)i * width;
            wgts[k] = weight[j] * half;
        }
}

}  // namespace gimic
