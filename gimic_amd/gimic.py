"""pygimic-facing `Gimic` class on the library's drop-in C entry points.

Mirrors the reference's Cython wrapper (src/libgimic/gimic.pyx:6-58):
Gimic(mol, xdens) calls gimic_init; jtensor, jvector and set_property keep
their names, arguments and type checks.  jtensor returns the nine tensor
elements (jt[m + 3b] = J(m, b)) as a length-9 array.  The batched methods
jtensors/jvectors evaluate many points in one call through the handle API.
"""
import numpy as np

from . import lib


class Gimic:
    def __init__(self, mol, xdens):
        self._lib = lib.load()
        self._lib.gimic_init(str(mol).encode(), str(xdens).encode())
        err = self._lib.gimic_last_error().decode()
        if err:
            raise lib.GimicError(err)
        self._magnet = np.zeros(3)
        self._spin = "total"
        self._ctx = None
        self._mol, self._xdens = str(mol), str(xdens)
        self._uhf = False
        self._thrs = 1e-6

    # -- gimic.pyx interface ------------------------------------------------
    def jtensor(self, r):
        cr = np.ascontiguousarray(np.asarray(r, dtype=np.float64)[:3])
        ct = np.zeros(9)
        self._lib.gimic_calc_jtensor(lib.ptr(cr), lib.ptr(ct))
        return ct

    def jvector(self, r):
        cr = np.ascontiguousarray(np.asarray(r, dtype=np.float64)[:3])
        cv = np.zeros(3)
        self._lib.gimic_calc_jvector(lib.ptr(cr), lib.ptr(cv))
        return [float(x) for x in cv]

    def set_property(self, prop, val):
        getattr(self, "set_" + prop)(val)

    def set_uhf(self, onoff):
        if not isinstance(onoff, int):
            raise TypeError
        import ctypes
        v = ctypes.c_int(onoff)
        self._lib.gimic_set_uhf(ctypes.byref(v))
        self._uhf = bool(onoff)

    def set_magnet(self, b):
        self._magnet = np.ascontiguousarray(np.asarray(b, dtype=np.float64)[:3])
        self._lib.gimic_set_magnet(lib.ptr(self._magnet))

    def set_spin(self, spin):
        if not isinstance(spin, str):
            raise TypeError
        self._lib.gimic_set_spin(spin.encode())
        self._spin = spin

    def set_screening(self, thrs):
        if not isinstance(thrs, float):
            raise TypeError
        import ctypes
        v = ctypes.c_double(thrs)
        self._lib.gimic_set_screening(ctypes.byref(v))
        self._thrs = thrs

    # -- batched extensions -----------------------------------------------------
    def _context(self):
        if self._ctx is None:
            self._ctx = lib.Context.from_files(self._mol, self._xdens, uhf=self._uhf,
                                               screening=self._thrs)
        return self._ctx

    def jtensors(self, points, precision="fp64"):
        """(N,3) -> (N,9) in one batched call"""
        return self._context().jtensor(points, spin=self._spin, precision=precision)

    def jvectors(self, points, precision="fp64"):
        J = self.jtensors(points, precision)
        return np.einsum("pbm,b->pm", J.reshape(-1, 3, 3), self._magnet)

    def close(self):
        if self._ctx is not None:
            self._ctx.close()
            self._ctx = None
        self._lib.gimic_finalize()
