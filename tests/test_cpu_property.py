"""CPU: the magnetic-property quadrature restatement (oracle/property_ref.py,
jfield.f90:584-929) pinned by closed forms.

For one normalised s Gaussian with no perturbed density the oracle's J is the
Larmor term, and get_property's sums must reduce to Lamb's diamagnetic
shielding (1e6 / 3c^2 * int rho / r, positive everywhere) and the Langevin
magnetizability (-1/6 int rho r^2, negative everywhere) on every diagonal
element.  A sign or factor slip in the restatement, or in the tensor layout
it reads, moves the result away from them.
"""
import numpy as np
import pytest

from oracle import molio, oracle, property_ref


def _one_s(alpha, R):
    b = molio.Basis()
    b.symbols, b.charges, b.coords = ["X"], [1.0], [list(R)]
    b.add_shell(0, 0, [alpha], [1.0])
    return b


@pytest.fixture(scope="module")
def s_atom():
    alpha, dval, R = 0.7, 2.0, np.array([0.3, -0.2, 0.1])
    b = _one_s(alpha, R)
    dens = np.zeros((1, 4, 1, 1))
    dens[0, 0] = dval
    pts, w = property_ref.spherical_grid(R, rmax=8.0 / np.sqrt(alpha))
    jt = oracle.jtensor(b.arrays(), dens, pts, screening=None)
    return alpha, dval, R, pts, w, jt


def test_s_shell_gives_lamb_shielding_and_langevin_magnetizability(s_atom):
    alpha, dval, R, pts, w, jt = s_atom
    sigma, chi = property_ref.s_shell_closed_forms(alpha, dval)
    sizes = [len(w) // 2, len(w) - len(w) // 2]  # two partitions: sums must add up
    shield, magnet = property_ref.properties(jt, pts, w, sizes, [R])
    tot_s = shield[0].sum(axis=0)
    tot_m = magnet.sum(axis=0)
    assert np.allclose(tot_s[:3], sigma, rtol=1e-9), (tot_s, sigma)
    assert np.allclose(tot_m[:3], chi, rtol=1e-9), (tot_m, chi)
    # about the nucleus the shielding integrand is positive everywhere; the
    # magnetizability integrand uses r about the origin, so only its parts' sum is fixed
    assert tot_s[4] == 0.0 and np.isclose(tot_s[3], 3 * sigma, rtol=1e-9)
    assert np.isclose(tot_m[3] + tot_m[4], 3 * chi, rtol=1e-9)
    assert sigma > 0 and chi < 0


def test_planted_sign_or_layout_slips_fail_the_closed_forms(s_atom):
    """mistakes planted in the oracle's inputs and arithmetic, on the CPU"""
    alpha, dval, R, pts, w, jt = s_atom
    sigma, chi = property_ref.s_shell_closed_forms(alpha, dval)
    # j = +J e_b instead of -J e_b
    s, m = property_ref.properties(-jt, pts, w, [len(w)], [R])
    assert not np.allclose(s[0, 0, :3], sigma, rtol=1e-3)
    # tensor read transposed (J(b, m) for J(m, b)): the Larmor term is antisymmetric
    jt_t = jt.reshape(-1, 3, 3).transpose(0, 2, 1).reshape(-1, 9)
    s, m = property_ref.properties(jt_t, pts, w, [len(w)], [R])
    assert not np.allclose(m[0, :3], chi, rtol=1e-3)
    # shielding taken about the origin instead of the nucleus
    s, m = property_ref.properties(jt, pts, w, [len(w)], [np.zeros(3)])
    assert not np.allclose(s[0, 0, :3], sigma, rtol=1e-3)


def test_magnetizability_does_not_depend_on_where_the_atom_sits():
    """the closed-shell current integrates to zero, so the r x j sum with the
    gauge origin at 0 is the same for the atom at the origin or displaced"""
    alpha, dval = 1.3, 2.0
    vals = []
    for R in (np.zeros(3), np.array([1.7, -0.4, 2.2])):
        b = _one_s(alpha, R)
        dens = np.zeros((1, 4, 1, 1))
        dens[0, 0] = dval
        pts, w = property_ref.spherical_grid(R, rmax=8.0 / np.sqrt(alpha), nr=48)
        jt = oracle.jtensor(b.arrays(), dens, pts, screening=None)
        m = property_ref.properties(jt, pts, w, [len(w)], [R])[1][0]
        if not R.any():  # atom at the gauge origin: every point's integrand is negative
            assert m[3] == 0.0
        vals.append(m[:3])
    assert np.allclose(vals[0], vals[1], rtol=1e-9)
    assert np.allclose(vals[0], property_ref.s_shell_closed_forms(alpha, dval)[1], rtol=1e-9)
