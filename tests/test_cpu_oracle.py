"""Pinning the CPU oracle (oracle/) against the reference and the mathematics.

* golden vectors from the flang-built reference (tests/golden/ring12_ref.npz)
  and the reference's own fixtures (c4h4 read-grid jvec.vtu)
* the flang-built reference itself where present (oracle/_ref)
* closed forms and invariants that a dropped term, wrong sign or swapped
  index would break
"""
import math

import numpy as np
import pytest

from conftest import golden, have_reference_build, rel_err, write_fixture
from oracle import molio, oracle, sph


def ring12_oracle_inputs():
    from gimic_amd import synth
    s = synth.ring12()
    b = molio.Basis()
    b.symbols, b.charges, b.coords = s.symbols, s.charges, s.coords
    b.shell_atom, b.shell_l, b.shell_nprim, b.shell_poff = (s.shell_atom, s.shell_l,
                                                             s.shell_nprim, s.shell_poff)
    b.exps, b.coefs = s.exps, s.coefs
    return s, b


def test_oracle_matches_reference_golden_ring12():
    g = golden("ring12_ref.npz")
    s, b = ring12_oracle_inputs()
    for flags in [(1, 1, 1), (0, 1, 1), (1, 0, 1), (1, 1, 0)]:
        J = oracle.jtensor(b.arrays(), g["dens_cart"], g["points"], screening=1e-8,
                           giao=flags[0], diamag=flags[1], paramag=flags[2])
        ref = g["jt_%d%d%d" % flags]
        assert np.abs(J - ref).max() <= 1e-14 * np.abs(ref).max(), flags
    J = oracle.jtensor(b.arrays(), g["dens_cart"], g["points"], screening=None)
    assert np.abs(J - g["jt_noscreen"]).max() <= 1e-14 * np.abs(g["jt_noscreen"]).max()


def test_oracle_matches_readgrid_fixture(tmp_path):
    # test/c4h4/read-grid: B = (0,0,-1), jvec = J B = -J(:,3), printed 3e20.10
    d = golden("c4h4.npz")
    mol, xd = write_fixture(str(tmp_path), "c4h4.npz")
    b = molio.read_mol(mol)
    assert b.turbomole and b.nbf == 168
    dens = molio.read_xdens(xd, b)
    pts = d["readgrid_points"][::7]
    ref = d["readgrid_jvec"][::7]
    J = oracle.jtensor(b.arrays(), dens, pts, screening=1e-8)
    jv = -J[:, 6:9]
    scale = np.abs(ref).max()
    assert np.all(np.abs(jv - ref) <= 5e-9 * np.abs(ref) + 1e-10 * scale)


@pytest.mark.skipif(not have_reference_build(), reason="reference tree or oracle/_ref absent")
@pytest.mark.parametrize("name,uhf", [("c4h4", False), ("open-shell", True)])
def test_oracle_bit_exact_with_flang_reference(name, uhf):
    from oracle import refrun
    d = f"/root/reference/test/{name}/"
    b = molio.read_mol(d + "MOL")
    dens = molio.read_xdens(d + "XDENS", b, uhf=uhf)
    pts = np.random.default_rng(3).uniform(-6, 6, (64, 3))
    for spin in (["total", "alpha", "beta", "spindens"] if uhf else ["total"]):
        Jr = refrun.ref_jtensor(d + "MOL", pts, xdens=d + "XDENS", uhf=uhf, spin=spin)
        Jo = oracle.jtensor(b.arrays(), dens, pts, spin=spin)
        assert np.array_equal(Jr, Jo), spin


# ---------------------------------------------------------------- invariants

def one_center(l=0, alpha=0.7, R=(0.3, -0.2, 0.1)):
    b = molio.Basis()
    b.symbols, b.charges, b.coords = ["X"], [1.0], [list(R)]
    b.add_shell(0, l, [alpha], [1.0])
    return b


def test_single_s_function_is_larmor_term():
    # one s function, DB = 0: J = 1/2 rho eps (r - R)  (jtensor.F90:225-235)
    alpha, R = 0.7, np.array([0.3, -0.2, 0.1])
    b = one_center(0, alpha, R)
    dens = np.zeros((1, 4, 1, 1))
    dens[0, 0] = 2.0
    r = np.array([[0.9, 0.4, -0.5]])
    J = oracle.jtensor(b.arrays(), dens, r, screening=None)[0].reshape(3, 3, order="F")
    N = (2 * alpha / math.pi) ** 0.75
    rho = 2.0 * (N * math.exp(-alpha * np.sum((r[0] - R) ** 2))) ** 2
    eps = np.zeros((3, 3, 3))
    eps[0, 1, 2] = eps[1, 2, 0] = eps[2, 0, 1] = 1
    eps[0, 2, 1] = eps[2, 1, 0] = eps[1, 0, 2] = -1
    # GIAO: one centre => only the gauge-origin-free part survives
    expect = 0.5 * rho * np.einsum("mbc,c->mb", eps, r[0] - R)
    assert np.allclose(J, expect, rtol=1e-13, atol=1e-16)


def test_normalisation_of_s_and_p_is_unit():
    # analytic overlap of contracted Gaussians with the normalised coefficients
    rng = np.random.default_rng(5)
    for l in (0, 1):
        b = molio.Basis()
        b.symbols, b.charges, b.coords = ["X"], [1.0], [[0.0, 0.0, 0.0]]
        ex = np.exp(rng.uniform(-2, 6, 4))
        b.add_shell(0, l, ex, rng.standard_normal(4))
        ncc, _ = oracle.normalise(b.arrays(), screening=None)
        S = 0.0
        for i in range(4):
            for j in range(4):
                p = ex[i] + ex[j]
                S += ncc[i] * ncc[j] * (math.pi / p) ** 1.5 * (1.0 / (2 * p)) ** l
        assert abs(S - 1.0) < 1e-12


def test_basis_gradient_matches_finite_differences():
    s, b = ring12_oracle_inputs()
    ba = b.arrays()
    r = np.array([0.37, -0.81, 0.42])
    _, grad = oracle.basis_values(ba, r, screening=None)
    h = 1e-5
    for ax in range(3):
        e = np.zeros(3)
        e[ax] = h
        fp, _ = oracle.basis_values(ba, r + e, screening=None)
        fm, _ = oracle.basis_values(ba, r - e, screening=None)
        fd = (fp - fm) / (2 * h)
        assert np.allclose(grad[ax], fd, rtol=1e-6, atol=1e-7 * np.abs(fd).max())


def test_translation_covariance_pins_giao_terms():
    # J(r+t; R+t, D, DB + dDB) = J(r; R, D, DB) with
    # dDB_b(mu,nu) = (t x (R_mu - R_nu))_b D(mu,nu)
    s, b = ring12_oracle_inputs()
    g = golden("ring12_ref.npz")
    D = g["dens_cart"][0]
    pts = g["points"][:40]
    t = np.array([1.3, -0.7, 2.1])
    ba = b.arrays()
    J0 = oracle.jtensor(ba, D[None], pts, screening=None)
    b2 = molio.Basis()
    b2.__dict__.update({k: list(v) if isinstance(v, list) else v for k, v in b.__dict__.items()})
    b2.coords = [list(np.asarray(c) + t) for c in b.coords]
    Rmu = np.asarray(b.coords)[np.repeat(np.asarray(b.shell_atom),
                                         [(l + 1) * (l + 2) // 2 for l in b.shell_l])]
    dR = Rmu[:, None, :] - Rmu[None, :, :]
    shift = np.cross(t[None, None, :], dR)  # (n,n,3)
    D2 = D.copy()
    for k in range(3):
        D2[1 + k] = D[1 + k] + shift[:, :, k] * D[0]
    J1 = oracle.jtensor(b2.arrays(), D2[None], pts + t, screening=None)
    assert rel_err(J1, J0).max() < 1e-9
    # dropping the GIAO correction breaks it (the test has teeth)
    J_bad = oracle.jtensor(b2.arrays(), D[None], pts + t, screening=None)
    assert rel_err(J_bad, J0).max() > 1e-3


def test_rotation_covariance_s_only_basis():
    rng = np.random.default_rng(11)
    b = molio.Basis()
    b.symbols, b.charges = ["X"] * 3, [1.0] * 3
    b.coords = rng.uniform(-1.5, 1.5, (3, 3)).tolist()
    for a in range(3):
        for _ in range(2):
            b.add_shell(a, 0, np.exp(rng.uniform(-1, 1, 2)), rng.standard_normal(2))
    n = b.nbf
    M = rng.standard_normal((n, n))
    D = np.zeros((4, n, n))
    D[0] = M @ M.T
    for k in range(3):
        A = rng.standard_normal((n, n))
        D[1 + k] = A - A.T
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    pts = rng.uniform(-2, 2, (20, 3))
    J = oracle.jtensor(b.arrays(), D[None], pts, screening=None).reshape(-1, 3, 3).transpose(0, 2, 1)
    b2 = molio.Basis()
    b2.__dict__.update(b.__dict__)
    b2.coords = (np.asarray(b.coords) @ Q.T).tolist()
    D2 = D.copy()
    for k in range(3):
        D2[1 + k] = sum(Q[k, j] * D[1 + j] for j in range(3))
    J2 = oracle.jtensor(b2.arrays(), D2[None], pts @ Q.T, screening=None)
    J2 = J2.reshape(-1, 3, 3).transpose(0, 2, 1)
    expect = np.einsum("ij,pjk,lk->pil", Q, J, Q)
    assert np.allclose(J2, expect, rtol=1e-10, atol=1e-12 * np.abs(J).max())


def test_linearity_spin_algebra_and_toggles(tmp_path):
    mol, xd = write_fixture(str(tmp_path), "openshell.npz")
    b = molio.read_mol(mol)
    dens = molio.read_xdens(xd, b, uhf=True)
    pts = np.random.default_rng(2).uniform(-5, 5, (24, 3))
    ba = b.arrays()
    Ja = oracle.jtensor(ba, dens, pts, spin="alpha")
    Jb = oracle.jtensor(ba, dens, pts, spin="beta")
    assert np.allclose(oracle.jtensor(ba, dens, pts, spin="total"), Ja + Jb, rtol=1e-13, atol=0)
    assert np.allclose(oracle.jtensor(ba, dens, pts, spin="spindens"), Ja - Jb, rtol=1e-13,
                       atol=1e-18)
    # J(D) via the alpha set equals J of the summed set given as closed shell
    Jsum = oracle.jtensor(ba, (dens[0] + dens[1])[None], pts)
    assert rel_err(Jsum, Ja + Jb).max() < 1e-12
    # toggles: para(DB) + giao(D) + dia(D) decomposition
    full = oracle.jtensor(ba, dens[:1], pts)
    nodia = oracle.jtensor(ba, dens[:1], pts, diamag=False)
    nopara = oracle.jtensor(ba, dens[:1], pts, paramag=False)
    dia_only = full - nodia
    assert np.allclose(nopara, dia_only, rtol=1e-12, atol=1e-14 * np.abs(full).max())
    dzero = dens[:1].copy()
    dzero[0, 1:] = 0.0
    nogiao_noDB = oracle.jtensor(ba, dzero, pts, giao=False, diamag=False)
    assert np.abs(nogiao_noDB).max() == 0.0


def test_screening_radius_is_reference_step_search():
    b = molio.Basis()
    b.symbols, b.charges, b.coords = ["X"], [1.0], [[0, 0, 0]]
    for l, a in [(0, 0.1), (1, 2.0), (2, 33.0), (3, 1e4)]:
        b.add_shell(0, l, [a, 3 * a], [1.0, 0.5])
    _, rad = oracle.normalise(b.arrays(), screening=1e-8)
    for (l, a), t in zip([(0, 0.1), (1, 2.0), (2, 33.0), (3, 1e4)], rad):
        k = round(t / 0.25)
        assert abs(t - 0.25 * k) < 1e-12
        assert t ** l * math.exp(-a * t * t) <= 1e-8
        tp = t - 0.25
        assert tp <= 0 or tp ** l * math.exp(-a * tp * tp) > 1e-8


def test_spherical_transform_oracle_properties():
    for l in range(2, 6):
        T = sph.transform(l)
        M = sph.cart_metric(l)
        assert np.allclose(T @ M @ T.T, np.eye(2 * l + 1), atol=1e-13)
        # rows are harmonic polynomials: Laplacian vanishes
        comp = sph.lex_components(l)
        for row in T:
            lap = {}
            for c, (a, bb, cc) in zip(row, comp):
                for ax, e in enumerate((a, bb, cc)):
                    if e >= 2:
                        k = [a, bb, cc]
                        k[ax] -= 2
                        lap[tuple(k)] = lap.get(tuple(k), 0.0) + c * e * (e - 1)
            assert max(abs(v) for v in lap.values()) < 1e-12
