"""CPU restatement of GIMIC's magnetic-property quadrature (test infrastructure).

Follows get_property (src/fgimic/jfield.f90:584-929): from the J tensors on
an external quadrature grid (points, weights, per-atom point blocks), the
diagonal of the nuclear shielding tensor of every nucleus (:691-745) and of
the magnetizability (:797-837).  For field direction b the current is
j_b = J (-e_b) (:707-727); with d = r - R_k and f = -1 / (|d|^3 c^2):
    shielding        s_b = 1e6 f (d x j_b)_b      (ppm)
    magnetizability  s_b = 1/2 (r x j_b)_b        (au, gauge origin 0)
Per point block (atom contribution) the weighted sums of s_x, s_y, s_z and
the positive and negative parts of the weighted sum of s_x + s_y + s_z.

Only tests/ use this module; the library computes the same sums on the GPU.
"""
import numpy as np

C_AU = 137.0359998  # jfield.f90:706
AU2SI_MAG = 7.89104e-29  # J/T^2 per au, jfield.f90:606


def _sums(s, w, offs):
    """per block: w-weighted sums of s[:, 0..2], positive and negative parts of w * sum"""
    pd = s.sum(axis=1)
    out = np.zeros((len(offs) - 1, 5))
    for p in range(len(offs) - 1):
        sl = slice(offs[p], offs[p + 1])
        ws = w[sl]
        out[p, :3] = (ws[:, None] * s[sl]).sum(axis=0)
        v = pd[sl] * ws
        out[p, 3] = v[v >= 0.0].sum()
        out[p, 4] = v[v < 0.0].sum()
    return out


def _currents(jt):
    """j[b] = J (-e_b) for b = x, y, z: (3, N, 3); J[m + 3b] is element (m, b)"""
    J = np.asarray(jt, dtype=np.float64).reshape(-1, 3, 3)  # [p, b, m]
    return np.stack([-J[:, b, :] for b in range(3)])


def _cross_diag(d, j):
    """(d x j_b)_b for b = x, y, z: (N, 3)"""
    out = np.empty((d.shape[0], 3))
    for b in range(3):
        b1, b2 = (b + 1) % 3, (b + 2) % 3
        out[:, b] = d[:, b1] * j[b][:, b2] - d[:, b2] * j[b][:, b1]
    return out


def properties(jt, points, weights, part_sizes, nuclei):
    """shield (K, P, 5) and magnet (P, 5), as gimic_ctx_properties"""
    r = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    w = np.asarray(weights, dtype=np.float64).reshape(-1)
    offs = np.concatenate([[0], np.cumsum(np.asarray(part_sizes, dtype=np.int64))])
    j = _currents(jt)
    nuc = np.asarray(nuclei, dtype=np.float64).reshape(-1, 3)
    shield = np.zeros((len(nuc), len(offs) - 1, 5))
    for k, R in enumerate(nuc):
        d = r - R
        f = -1.0 / (np.sum(d * d, axis=1) ** 1.5) / C_AU ** 2
        shield[k] = _sums(1.0e6 * f[:, None] * _cross_diag(d, j), w, offs)
    magnet = _sums(0.5 * _cross_diag(r, j), w, offs)
    return shield, magnet


def s_shell_closed_forms(alpha, dval):
    """One normalised s Gaussian exp(-alpha r^2) with density coefficient D
    and no perturbed density: J is the Larmor term 1/2 eps_mbc rho r_c, and
    get_property's sums reduce to Lamb's diamagnetic shielding and the
    Langevin magnetizability, per diagonal element:
        sigma = 1e6 / (3 c^2) * int rho / r = 1e6 / (3 c^2) * D N^2 pi / alpha
        chi   = -1/4 int rho (y^2 + z^2) = -1/6 int rho r^2 = -D / (8 alpha)
    with N^2 = (2 alpha / pi)^(3/2)."""
    n2 = (2.0 * alpha / np.pi) ** 1.5
    sigma = 1.0e6 / (3.0 * C_AU ** 2) * dval * n2 * np.pi / alpha
    chi = -dval / (8.0 * alpha)
    return sigma, chi


def spherical_grid(center, rmax, nr=64, nt=16, nphi=32):
    """product quadrature on a ball: Gauss-Legendre in r on [0, rmax] and in
    cos(theta), trapezoid in phi; exact for the angular degree the closed
    forms need"""
    xr, wr = np.polynomial.legendre.leggauss(nr)
    rr = 0.5 * rmax * (xr + 1.0)
    wr = 0.5 * rmax * wr * rr ** 2
    ct, wt = np.polynomial.legendre.leggauss(nt)
    ph = 2.0 * np.pi * np.arange(nphi) / nphi
    wp = np.full(nphi, 2.0 * np.pi / nphi)
    R, CT, PH = np.meshgrid(rr, ct, ph, indexing="ij")
    W = wr[:, None, None] * wt[None, :, None] * wp[None, None, :]
    st = np.sqrt(1.0 - CT ** 2)
    pts = np.stack([R * st * np.cos(PH), R * st * np.sin(PH), R * CT], axis=-1).reshape(-1, 3)
    return pts + np.asarray(center, dtype=np.float64), W.reshape(-1)
