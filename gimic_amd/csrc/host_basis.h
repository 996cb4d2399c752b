// Host-side basis and density handling for the GIMIC current-density library.
// Reference conventions: src/libgimic/{intgrl,basis,gtodefs,dens,reorder}.f90.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace gimic {

constexpr int kMaxL = 5;  // globals.f90:30 MAX_L

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// One segmented contraction ("shell") in reference function order.
struct Shell {
    int atom;    // 0-based atom index
    int l;       // angular momentum
    int nprim;   // primitives
    int poff;    // offset into exps/coefs/ncc
    int fcart;   // first cartesian function index
    int fout;    // first output function index (cartesian or spherical)
    double radius;  // screening radius (setup_screening), 1e10 when off
};

struct Basis {
    std::vector<double> coords;      // 3*natoms, bohr
    std::vector<double> charges;
    std::vector<std::string> symbols;
    std::vector<Shell> shells;
    std::vector<double> exps, coefs, ncc;
    bool turbomole = false;
    bool spherical = false;
    int nbf_cart = 0;
    int nbf = 0;  // output basis size (== nbf_cart unless spherical)
    int natoms() const { return (int)charges.size(); }
};

// Parse a MOL/INTGRL file (intgrl.f90:20-264). Throws std::runtime_error.
Basis read_mol(const std::string &path);

// Build from arrays (shells must be atom-major, atoms 0..natoms-1).
Basis make_basis(int natoms, const double *coords, int nshell, const int *shell_atom,
                 const int *shell_l, const int *shell_nprim, const int *shell_poff,
                 const double *exps, const double *coefs, bool turbomole, bool spherical);

// Contraction normalisation (basis.f90:164-191) and screening radii
// (basis.f90:90-112); thrs <= 0 switches screening off (radius 1e10).
void finalize_basis(Basis &b, double screen_thrs);

// Cartesian exponent table for l (gtodefs.f90:86-123), row-major ncart x 3.
void cart_components(int l, bool turbomole, int out[][3]);

// Real solid-harmonic transform, (2l+1) x ncart(l) row-major, m = -l..l,
// rows normalised under the reference's cartesian normalisation.
std::vector<double> sph_transform(int l);

// XDENS reader (dens.f90:56-135): nset*4 blocks of n*n column-major values,
// UHF perturbed blocks halved, TURBOMOLE l-sorted order permuted back to
// atom-major (reorder.f90:54-96, dens.f90:210-234).
std::vector<double> read_xdens(const std::string &path, const Basis &b, int nset, bool uhf);

}  // namespace gimic
