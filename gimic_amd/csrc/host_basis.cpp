// Host-side basis and density handling (see host_basis.h for the reference
// files each routine follows).
#include "host_basis.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace gimic {

namespace {

// TURBOMOLE cartesian orders, gtodefs.f90:108-123.
const int kTmD[6][3] = {{2,0,0},{0,2,0},{0,0,2},{1,1,0},{1,0,1},{0,1,1}};
const int kTmF[10][3] = {{3,0,0},{0,3,0},{0,0,3},{2,1,0},{2,0,1},
                         {1,2,0},{0,2,1},{1,0,2},{0,1,2},{1,1,1}};
const int kTmG[15][3] = {{4,0,0},{0,4,0},{0,0,4},{3,1,0},{3,0,1},{1,3,0},{0,3,1},
                         {1,0,3},{0,1,3},{2,2,0},{2,0,2},{0,2,2},{2,1,1},{1,2,1},{1,1,2}};
const int kTmH[21][3] = {{5,0,0},{0,5,0},{0,0,5},{4,1,0},{4,0,1},{1,4,0},{0,4,1},
                         {1,0,4},{0,1,4},{3,2,0},{3,0,2},{2,3,0},{0,3,2},{2,0,3},
                         {0,2,3},{3,1,1},{1,3,1},{1,1,3},{2,2,1},{2,1,2},{1,2,2}};

double fnum(std::string tok) {
    for (auto &c : tok)
        if (c == 'D' || c == 'd') c = 'E';
    char *end = nullptr;
    double v = std::strtod(tok.c_str(), &end);
    if (end == tok.c_str()) throw std::runtime_error("bad number '" + tok + "'");
    return v;
}

// Fortran list-directed reading: every read starts on a new record and
// consumes as many records as it needs.
struct Records {
    std::vector<std::string> lines;
    size_t i = 0;
    std::string line() {
        if (i >= lines.size()) throw std::runtime_error("unexpected end of file");
        return lines[i++];
    }
    std::vector<std::string> values(size_t n) {
        std::vector<std::string> v;
        while (v.size() < n) {
            std::string s = line();
            for (auto &c : s)
                if (c == ',') c = ' ';
            std::istringstream ss(s);
            std::string t;
            while (ss >> t) v.push_back(t);
        }
        v.resize(n);
        return v;
    }
};

double dfact2(int n) {  // (n)!! with (-1)!! = 1
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
}

double binom(int n, int k) {
    if (k < 0 || k > n) return 0.0;
    double r = 1.0;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

}  // namespace

void cart_components(int l, bool turbomole, int out[][3]) {
    if (turbomole && l >= 2) {
        const int(*t)[3] = l == 2 ? kTmD : l == 3 ? kTmF : l == 4 ? kTmG : kTmH;
        for (int i = 0; i < ncart(l); i++)
            for (int k = 0; k < 3; k++) out[i][k] = t[i][k];
        return;
    }
    int n = 0;  // lexicographic order of gtodefs.f90:86-106
    for (int a = l; a >= 0; a--)
        for (int b = l - a; b >= 0; b--) {
            out[n][0] = a;
            out[n][1] = b;
            out[n][2] = l - a - b;
            n++;
        }
}

std::vector<double> sph_transform(int l) {
    // Real solid harmonics as polynomials in x,y,z (Helgaker, Jorgensen,
    // Olsen, "Molecular Electronic-Structure Theory", eqs. 6.4.47-6.4.50),
    // then each row scaled to unit norm under the cartesian metric of the
    // reference normalisation: <a b c|a' b' c'> = (a+a'-1)!!(b+b'-1)!!(c+c'-1)!!.
    const int nc = ncart(l), ns = 2 * l + 1;
    int comp[21][3];
    cart_components(l, false, comp);
    auto cidx = [&](int a, int b, int c) {
        for (int i = 0; i < nc; i++)
            if (comp[i][0] == a && comp[i][1] == b && comp[i][2] == c) return i;
        return -1;
    };
    std::vector<double> T((size_t)ns * nc, 0.0);
    for (int m = -l; m <= l; m++) {
        const int am = std::abs(m);
        const double vm = m < 0 ? 0.5 : 0.0;
        double *row = &T[(size_t)(m + l) * nc];
        for (int t = 0; t <= (l - am) / 2; t++)
            for (int u = 0; u <= t; u++) {
                // v runs over vm, vm+1, ... up to floor(am/2 - vm) + vm
                const int vmax2 = 2 * ((int)std::floor(am / 2.0 - vm)) + (int)(2 * vm);
                for (int v2 = (int)(2 * vm); v2 <= vmax2; v2 += 2) {
                    const double v = v2 / 2.0;
                    const double sgn = ((int)std::lround(t + v - vm)) % 2 ? -1.0 : 1.0;
                    const double c = sgn * std::pow(0.25, t) * binom(l, t) * binom(l - t, am + t) *
                                     binom(t, u) * binom(am, v2);
                    const int px = 2 * t + am - 2 * u - v2, py = 2 * u + v2, pz = l - 2 * t - am;
                    const int k = cidx(px, py, pz);
                    if (k < 0) throw std::logic_error("sph_transform index");
                    row[k] += c;
                }
            }
        double nrm = 0.0;
        for (int i = 0; i < nc; i++)
            for (int j = 0; j < nc; j++) {
                const int A = comp[i][0] + comp[j][0], B = comp[i][1] + comp[j][1],
                          C = comp[i][2] + comp[j][2];
                if ((A | B | C) & 1) continue;
                nrm += row[i] * row[j] * dfact2(A - 1) * dfact2(B - 1) * dfact2(C - 1);
            }
        nrm = 1.0 / std::sqrt(nrm);
        for (int i = 0; i < nc; i++) row[i] *= nrm;
    }
    return T;
}

static void layout(Basis &b) {
    int fc = 0, fo = 0;
    int prev_atom = 0;
    for (auto &s : b.shells) {
        if (s.atom < prev_atom || s.atom >= b.natoms())
            throw std::runtime_error("shells must be atom-major with valid atom indices");
        if (s.l < 0 || s.l > kMaxL || s.nprim < 1)
            throw std::runtime_error("invalid shell (l or nprim)");
        prev_atom = s.atom;
        s.fcart = fc;
        s.fout = fo;
        fc += ncart(s.l);
        fo += b.spherical ? 2 * s.l + 1 : ncart(s.l);
    }
    b.nbf_cart = fc;
    b.nbf = fo;
}

Basis read_mol(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open basis file '" + path + "'");
    Records rec;
    for (std::string s; std::getline(f, s);) {
        if (!s.empty() && s.back() == '\r') s.pop_back();
        rec.lines.push_back(s);
    }
    Basis b;
    if (rec.line().compare(0, 6, "INTGRL") != 0)
        throw std::runtime_error("'" + path + "' does not look like an INTGRL file");
    b.turbomole = rec.line().compare(0, 9, "TURBOMOLE") == 0;
    rec.line();
    const int natoms = std::atoi(rec.values(1)[0].c_str());
    rec.line();
    if (natoms <= 0) throw std::runtime_error("invalid number of atoms");
    for (int a = 0; a < natoms; a++) {
        // charge, dummy, nshells, contractions per shell (list-directed)
        std::vector<std::string> hdr = rec.values(3);
        rec.i--;
        const int nsh = std::atoi(hdr[2].c_str());
        if (nsh < 1 || nsh > kMaxL + 1)
            throw std::runtime_error("largest allowed l-quantum number exceeded");
        hdr = rec.values(3 + nsh);
        b.charges.push_back(fnum(hdr[0]));
        const std::string sl = rec.line();
        b.symbols.push_back(sl.substr(0, std::min<size_t>(2, sl.size())));
        std::istringstream cs(sl.size() > 4 ? sl.substr(4) : std::string());
        std::string t;
        for (int k = 0; k < 3; k++) {
            if (!(cs >> t)) throw std::runtime_error("bad coordinate line: " + sl);
            b.coords.push_back(fnum(t));
        }
        for (int ish = 0; ish < nsh; ish++) {
            const int nblk = std::atoi(hdr[3 + ish].c_str());
            for (int j = 0; j < nblk; j++) {
                auto pn = rec.values(2);
                const int npf = std::atoi(pn[0].c_str()), ncf = std::atoi(pn[1].c_str());
                if (npf < 1 || ncf < 1) throw std::runtime_error("bad contraction header");
                std::vector<double> xp(npf), cc((size_t)npf * ncf);
                for (int p = 0; p < npf; p++) {
                    auto v = rec.values(1 + ncf);
                    xp[p] = fnum(v[0]);
                    for (int k = 0; k < ncf; k++) cc[(size_t)p * ncf + k] = fnum(v[1 + k]);
                }
                // a general contraction becomes ncf segmented ones
                // (intgrl.f90:172-216)
                for (int k = 0; k < ncf; k++) {
                    Shell s{a, ish, npf, (int)b.exps.size(), 0, 0, 1e10};
                    for (int p = 0; p < npf; p++) {
                        b.exps.push_back(xp[p]);
                        b.coefs.push_back(cc[(size_t)p * ncf + k]);
                    }
                    b.shells.push_back(s);
                }
            }
        }
    }
    layout(b);
    return b;
}

Basis make_basis(int natoms, const double *coords, int nshell, const int *shell_atom,
                 const int *shell_l, const int *shell_nprim, const int *shell_poff,
                 const double *exps, const double *coefs, bool turbomole, bool spherical) {
    if (natoms <= 0 || nshell <= 0) throw std::runtime_error("empty basis");
    Basis b;
    b.turbomole = turbomole;
    b.spherical = spherical;
    b.coords.assign(coords, coords + 3 * natoms);
    b.charges.assign(natoms, 0.0);
    b.symbols.assign(natoms, "X ");
    for (int s = 0; s < nshell; s++) {
        Shell sh{shell_atom[s], shell_l[s], shell_nprim[s], (int)b.exps.size(), 0, 0, 1e10};
        if (sh.nprim < 1) throw std::runtime_error("shell without primitives");
        for (int p = 0; p < sh.nprim; p++) {
            b.exps.push_back(exps[shell_poff[s] + p]);
            b.coefs.push_back(coefs[shell_poff[s] + p]);
        }
        b.shells.push_back(sh);
    }
    layout(b);
    return b;
}

void finalize_basis(Basis &b, double thrs) {
    const double PII = 3.141592653589793;  // globals.f90 PII
    b.ncc.assign(b.exps.size(), 0.0);
    for (auto &s : b.shells) {
        const double *xp = &b.exps[s.poff], *cc = &b.coefs[s.poff];
        // norm_ctr (basis.f90:164-191)
        const double j = 1.0 * (s.l + 1);
        double n = 0.0;
        for (int a = 0; a < s.nprim; a++)
            for (int c = 0; c <= a; c++) {
                double t = 2.0 * std::sqrt(xp[a] * xp[c]) / (xp[a] + xp[c]);
                t = cc[a] * cc[c] * std::pow(t, j + 0.5);
                n += (a != c) ? 2.0 * t : t;
            }
        if (!(n > 0.0)) throw std::runtime_error("contraction with zero norm");
        n = 1.0 / std::sqrt(n);
        for (int a = 0; a < s.nprim; a++)
            b.ncc[s.poff + a] = cc[a] * n * std::pow(4.0 * xp[a], 0.5 * j + 0.25) *
                                std::pow(0.5 / PII, 0.75);
        // setup_screening (basis.f90:90-112): first multiple of 0.25 bohr at
        // which t^l exp(-alpha_min t^2) drops to the threshold
        if (thrs <= 0.0) {
            s.radius = 1e10;
        } else {
            double amin = 1e15, x = 1e15, dist = 0.0;
            for (int a = 0; a < s.nprim; a++) amin = std::min(amin, xp[a]);
            while (x > thrs) {
                dist += 0.25;
                x = std::pow(dist, s.l) * std::exp(-amin * dist * dist);
            }
            s.radius = dist;
        }
    }
}

std::vector<double> read_xdens(const std::string &path, const Basis &b, int nset, bool uhf) {
    FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) throw std::runtime_error("density file '" + path + "' not found");
    std::fseek(fp, 0, SEEK_END);
    const long sz = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    std::string buf((size_t)sz, '\0');
    if (std::fread(&buf[0], 1, (size_t)sz, fp) != (size_t)sz) {
        std::fclose(fp);
        throw std::runtime_error("short read of '" + path + "'");
    }
    std::fclose(fp);
    const size_t n = (size_t)b.nbf_cart, need = (size_t)nset * 4 * n * n;
    std::vector<double> v(need);
    size_t k = 0;
    const char *p = buf.data(), *end = p + buf.size();
    // one value per record: the first token of each non-blank line
    while (p < end && k < need) {
        const char *eol = (const char *)std::memchr(p, '\n', (size_t)(end - p));
        if (!eol) eol = end;
        const char *q = p;
        while (q < eol && (*q == ' ' || *q == '\t' || *q == '\r')) q++;
        if (q < eol) {
            const char *te = q;
            while (te < eol && *te != ' ' && *te != '\t' && *te != '\r' && *te != ',') te++;
            v[k++] = fnum(std::string(q, te));
        }
        p = eol + 1;
    }
    if (k < need) throw std::runtime_error("density file '" + path + "' is too short");
    if (uhf)  // dens.f90:94-98
        for (int s = 0; s < nset; s++)
            for (int m = 1; m < 4; m++) {
                double *blk = &v[((size_t)s * 4 + m) * n * n];
                for (size_t i = 0; i < n * n; i++) blk[i] *= 0.5;
            }
    if (b.turbomole) {
        // file order: all l=0 functions (atom-major), then l=1, ...;
        // sv[q] = atom-major index of file position q (reorder.f90:54-96)
        std::vector<int> lvec;
        for (auto &s : b.shells)
            for (int c = 0; c < ncart(s.l); c++) lvec.push_back(s.l);
        std::vector<size_t> sv;
        for (int l = 0; l <= kMaxL; l++)
            for (size_t i = 0; i < lvec.size(); i++)
                if (lvec[i] == l) sv.push_back(i);
        std::vector<double> tmp(n * n);
        for (int blk = 0; blk < nset * 4; blk++) {
            double *d = &v[(size_t)blk * n * n];
            for (size_t j = 0; j < n; j++)
                for (size_t i = 0; i < n; i++) tmp[sv[i] + n * sv[j]] = d[i + n * j];
            std::memcpy(d, tmp.data(), sizeof(double) * n * n);
        }
    }
    return v;
}

}  // namespace gimic
