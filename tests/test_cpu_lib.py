"""Host-side checks of the C-ABI library that need no GPU: it loads, exports
every symbol the header declares, and its host-only routines (quadrature,
spherical transform) agree with the reference and the mathematics."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, golden
from gimic_amd import lib


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "gimic_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gimic_\w+|mkgausspoints)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = lib.load()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert set(lib.EXPORTS) == set(syms)


def test_quadrature_matches_reference_setup_gauss_data():
    g = golden("ring12_ref.npz")
    for kind in ("gauss", "lobatto"):
        for order in (3, 5, 7, 9, 11):
            p, w = lib.quadrature(0.0, 7.3, order, 4 * order, kind)
            # the same start values, steps and operation order: the same bits
            assert np.array_equal(p, g[f"{kind}_{order}_pts"]), (kind, order)
            assert np.array_equal(w, g[f"{kind}_{order}_wgt"]), (kind, order)


def test_quadrature_against_textbook_rules():
    for n in range(2, 14):
        x, w = np.polynomial.legendre.leggauss(n)
        p, wt = lib.quadrature(-1.0, 1.0, n, n, "gauss")
        # setup_gauss_data maps [-1,1] onto [0, b-a]
        assert np.allclose(p - 1.0, x, atol=1e-13) and np.allclose(wt, w, atol=1e-13)
    for n in range(3, 14):
        P = np.polynomial.legendre.Legendre.basis(n - 1)
        inner = np.sort(P.deriv().roots().real)
        x = np.concatenate([[-1.0], inner, [1.0]])
        w = 2.0 / (n * (n - 1) * P(x) ** 2)
        p, wt = lib.quadrature(-1.0, 1.0, n, n, "lobatto")
        assert np.allclose(np.sort(p - 1.0), x, atol=1e-12)
        assert np.allclose(wt[np.argsort(p)], w, atol=1e-12)


def test_composite_rules_integrate_polynomials_exactly():
    for kind, deg in (("gauss", lambda n: 2 * n - 1), ("lobatto", lambda n: 2 * n - 3)):
        for n in (3, 7, 9):
            p, w = lib.quadrature(0.0, 5.0, n, 3 * n, kind)
            k = deg(n)
            # exact per block: integrate x^k over [0,5]
            assert abs(np.sum(w * p ** k) - 5.0 ** (k + 1) / (k + 1)) < 1e-10 * 5.0 ** (k + 1)
    p, w = lib.quadrature(0.0, 3.0, 7, 1, "gauss")
    assert p[0] == 0.0 and w[0] == 1.0


def test_quadrature_rejects_bad_sizes():
    with pytest.raises(lib.GimicError):
        lib.quadrature(0.0, 1.0, 7, 10, "gauss")


def test_drop_in_gauss_points_symbol():
    L = lib.load()
    a, b = ctypes.c_double(0.0), ctypes.c_double(2.0)
    n, o = ctypes.c_int(6), ctypes.c_int(3)
    p = np.zeros(6)
    w = np.zeros(6)
    L.gimic_get_gauss_points(ctypes.byref(a), ctypes.byref(b), ctypes.byref(n), ctypes.byref(o),
                             lib.ptr(p), lib.ptr(w))
    q, v = lib.quadrature(0.0, 2.0, 3, 6, "gauss")
    assert np.array_equal(p, q) and np.array_equal(w, v)


def test_spherical_transform_matches_independent_recursion():
    from oracle import sph
    for l in range(0, 6):
        T = lib.sph_transform(l)
        if l >= 2:
            assert np.allclose(T, sph.transform(l), atol=1e-14)
        else:
            assert np.allclose(T, np.eye((l + 1) * (l + 2) // 2))


def test_spherical_transform_zeros_follow_parity():
    """The basis kernel skips, at compile time, every T_l entry whose
    monomial x^a y^b z^c has the wrong parity for row m: b odd iff m < 0,
    a + b of the parity of |m|, c of the parity of l - |m|.  Those entries
    must be exactly zero in the table the library uploads, and the rule
    must not be vacuous (rows keep at least one term, most entries go)."""
    for l in range(2, 6):
        T = np.asarray(lib.sph_transform(l))
        comps = [(a, b, l - a - b) for a in range(l, -1, -1) for b in range(l - a, -1, -1)]
        kept = 0
        for mi in range(2 * l + 1):
            m = mi - l
            yp = 1 if m < 0 else 0
            mask = [(a & 1) == ((abs(m) + yp) & 1) and (b & 1) == yp and (c & 1) == ((l - abs(m)) & 1)
                    for a, b, c in comps]
            assert all(T[mi, k] == 0.0 for k in range(len(comps)) if not mask[k]), (l, m)
            assert any(T[mi, k] != 0.0 for k in range(len(comps)) if mask[k]), (l, m)
            kept += sum(mask)
        assert kept < 0.4 * T.size, (l, kept, T.size)
