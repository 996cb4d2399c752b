"""Real solid harmonics from their recurrence (test infrastructure only).

Independent of the library's closed-form construction: polynomials are
built with the standard recursion for real solid harmonics (Helgaker,
Jorgensen, Olsen, "Molecular Electronic-Structure Theory", eqs. 6.4.70-73)
and normalised under the reference's cartesian normalisation, where
<a b c|a' b' c'> = (a+a'-1)!!(b+b'-1)!!(c+c'-1)!! (basis.f90:164-191 gives
unit norm to the components with all exponents <= 1).

Used to turn a spherical density family into the cartesian one the
reference consumes: D_cart = T^T D_sph T (J is bilinear in (phi, D)).
"""
import math
from collections import defaultdict

import numpy as np


def lex_components(l):
    return [(a, b, l - a - b) for a in range(l, -1, -1) for b in range(l - a, -1, -1)]


def _mul(p, q):
    out = defaultdict(float)
    for k1, v1 in p.items():
        for k2, v2 in q.items():
            out[(k1[0] + k2[0], k1[1] + k2[1], k1[2] + k2[2])] += v1 * v2
    return dict(out)


def _add(p, q, a=1.0, b=1.0):
    out = defaultdict(float)
    for k, v in p.items():
        out[k] += a * v
    for k, v in q.items():
        out[k] += b * v
    return dict(out)


X, Y, Z = {(1, 0, 0): 1.0}, {(0, 1, 0): 1.0}, {(0, 0, 1): 1.0}
R2 = {(2, 0, 0): 1.0, (0, 2, 0): 1.0, (0, 0, 2): 1.0}


def solid_harmonics(lmax):
    """S[l][m] as polynomial dicts, m = -l..l (Racah normalisation)."""
    S = {0: {0: {(0, 0, 0): 1.0}}}
    for l in range(lmax):
        S[l + 1] = {}
        f = math.sqrt((2.0 if l == 0 else 1.0) * (2 * l + 1) / (2 * l + 2))
        d = 0.0 if l == 0 else 1.0
        Sll, Slm = S[l][l], S[l][-l]
        S[l + 1][l + 1] = {k: f * v for k, v in _add(_mul(X, Sll), _mul(Y, Slm), 1.0, -d).items()}
        S[l + 1][-l - 1] = {k: f * v for k, v in _add(_mul(Y, Sll), _mul(X, Slm), 1.0, d).items()}
        for m in range(-l, l + 1):
            a = _mul(Z, S[l][m])
            b = _mul(R2, S[l - 1][m]) if (l - 1) >= abs(m) else {}
            c1 = (2 * l + 1) / math.sqrt((l + m + 1) * (l - m + 1))
            c2 = math.sqrt((l + m) * (l - m)) / math.sqrt((l + m + 1) * (l - m + 1))
            S[l + 1][m] = _add(a, b, c1, -c2)
    return S


def _df(n):
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def cart_metric(l):
    comp = lex_components(l)
    M = np.zeros((len(comp), len(comp)))
    for i, (a, b, c) in enumerate(comp):
        for j, (a2, b2, c2) in enumerate(comp):
            A, B, C = a + a2, b + b2, c + c2
            if A % 2 or B % 2 or C % 2:
                continue
            M[i, j] = _df(A - 1) * _df(B - 1) * _df(C - 1)
    return M


def transform(l):
    """T[(m+l), c] for m=-l..l over lexicographic cartesian components."""
    if l <= 1:
        return np.eye((l + 1) * (l + 2) // 2)
    S = solid_harmonics(l)[l]
    comp = lex_components(l)
    T = np.zeros((2 * l + 1, len(comp)))
    for m in range(-l, l + 1):
        for k, v in S[m].items():
            T[m + l, comp.index(k)] += v
    M = cart_metric(l)
    for i in range(T.shape[0]):
        T[i] /= math.sqrt(T[i] @ M @ T[i])
    return T


def basis_transform(shell_l):
    """block-diagonal T (nsph x ncart) for a list of shell angular momenta"""
    blocks = [transform(l) for l in shell_l]
    ns = sum(b.shape[0] for b in blocks)
    nc = sum(b.shape[1] for b in blocks)
    T = np.zeros((ns, nc))
    i = j = 0
    for b in blocks:
        T[i:i + b.shape[0], j:j + b.shape[1]] = b
        i += b.shape[0]
        j += b.shape[1]
    return T


def to_cartesian(dens, shell_l):
    """(nset,4,nsph,nsph) spherical family -> cartesian D_cart = T^T D T
    (T is block-diagonal by shell: the products skip its exact zeros)"""
    from scipy import sparse
    Ts = sparse.csr_matrix(basis_transform(shell_l))
    d = np.asarray(dens, dtype=np.float64)
    nc = Ts.shape[1]
    out = np.empty(d.shape[:2] + (nc, nc))
    for s in range(d.shape[0]):
        for k in range(d.shape[1]):
            x = np.asarray(Ts.T @ d[s, k])          # T^T D        (nc, nsph)
            out[s, k] = np.asarray(Ts.T @ x.T).T    # (T^T (T^T D)^T)^T = T^T D T
    return out
