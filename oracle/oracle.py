"""ctypes wrapper of the C restatement (liboracle.so). Test infrastructure only."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
SPIN = {"total": 0, "alpha": 1, "beta": 2, "spindens": 3}
_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "liboracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        L.oracle_jtensor.argtypes = [ctypes.c_int, P, ctypes.c_int, P, P, P, P, P, P,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_int, P,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int64, P, P]
        L.oracle_basis.argtypes = [ctypes.c_int, P, ctypes.c_int, P, P, P, P, P, P,
                                   ctypes.c_int, ctypes.c_double, P, P, P]
        L.oracle_normalise.argtypes = [ctypes.c_int, P, P, P, P, P, ctypes.c_double, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _basis_args(ba):
    c = ba["coords"]
    return [c.shape[0], _p(c), len(ba["shell_l"]), _p(ba["shell_atom"]), _p(ba["shell_l"]),
            _p(ba["shell_nprim"]), _p(ba["shell_poff"]), _p(ba["exps"]), _p(ba["coefs"]),
            int(ba["turbomole"])]


def jtensor(basis_arrays, dens, points, *, screening=1e-8, spin="total", giao=True,
            diamag=True, paramag=True):
    """J tensors (N,9), ct(m,b) column-major; dens is (nset,4,n,n) [mu,nu]."""
    from .molio import pack_dens
    ba = basis_arrays
    d = np.asarray(dens, dtype=np.float64)
    nset, n = d.shape[0], d.shape[-1]
    flat = pack_dens(d)
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    out = np.zeros((pts.shape[0], 9))
    rc = lib().oracle_jtensor(*_basis_args(ba), float(screening or -1.0), n, _p(flat), nset,
                              SPIN[spin], int(giao), int(diamag), int(paramag),
                              pts.shape[0], _p(pts), _p(out))
    if rc:
        raise RuntimeError(f"oracle_jtensor rc={rc}")
    return out


def basis_values(basis_arrays, r, *, screening=1e-8):
    ba = basis_arrays
    n = int(sum((l + 1) * (l + 2) // 2 for l in ba["shell_l"]))
    r = np.ascontiguousarray(np.asarray(r, dtype=np.float64))
    phi = np.zeros(n)
    grad = np.zeros(3 * n)
    rc = lib().oracle_basis(*_basis_args(ba), float(screening or -1.0), _p(r), _p(phi), _p(grad))
    if rc:
        raise RuntimeError(f"oracle_basis rc={rc}")
    return phi, grad.reshape(3, n)


def normalise(basis_arrays, screening=1e-8):
    ba = basis_arrays
    ncc = np.zeros(len(ba["exps"]))
    rad = np.zeros(len(ba["shell_l"]))
    rc = lib().oracle_normalise(len(ba["shell_l"]), _p(ba["shell_l"]), _p(ba["shell_nprim"]),
                                _p(ba["shell_poff"]), _p(ba["exps"]), _p(ba["coefs"]),
                                float(screening or -1.0), _p(ncc), _p(rad))
    if rc:
        raise RuntimeError(f"oracle_normalise rc={rc}")
    return ncc, rad
