"""Runner for the flang-built reference oracle (test infrastructure only).

`oracle/_ref/ref_driver` is this project's Fortran driver linked against the
reference libgimic modules compiled from their own sources (oracle/Makefile).
It evaluates the reference per-point tensor (jtensor.F90:66-123, `ctensor`)
at arbitrary points, so tests can compare any implementation with the
reference itself.  Only tests, smoke checks and the bench CPU baseline use it.
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
DRIVER = os.path.join(REF_DIR, "ref_driver")
GAUSS = os.path.join(REF_DIR, "ref_gauss")


def available():
    return os.path.exists(DRIVER) and os.path.exists(GAUSS)


def ref_jtensor(molfile, points, *, xdens=None, dens=None, uhf=False,
                screening=1e-8, giao=True, diamag=True, paramag=True,
                spin="total", threads=None, return_time=False):
    """Reference J tensors at `points` (N,3) -> (N,9), ct(m,b) column-major.

    Either `xdens` (a text XDENS path, read by the reference reader with its
    UHF 1/2 scaling and TURBOMOLE reorder) or `dens` (array of 4 or 8 n*n
    column-major matrices, passed verbatim) must be given.
    """
    if not available():
        raise RuntimeError("reference oracle not built (make -C oracle ref)")
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    with tempfile.TemporaryDirectory(prefix="gimic_ref_") as td:
        pfile = os.path.join(td, "pts.bin")
        ofile = os.path.join(td, "out.bin")
        with open(pfile, "wb") as f:
            np.array([pts.shape[0]], dtype=np.int32).tofile(f)
            pts.tofile(f)
        if xdens is not None:
            mode, dfile = "xdens", os.path.abspath(xdens)
        else:
            mode, dfile = "bin", os.path.join(td, "dens.bin")
            np.ascontiguousarray(np.asarray(dens, dtype=np.float64)).tofile(dfile)
        ctl = "\n".join([
            os.path.abspath(molfile), mode, dfile, str(int(bool(uhf))),
            repr(float(screening) if screening else -1.0),
            f"{int(giao)} {int(diamag)} {int(paramag)}", spin, pfile, ofile, ""])
        env = dict(os.environ)
        if threads is not None:
            env["OMP_NUM_THREADS"] = str(threads)
        # the reference writes mol.xyz into the working directory
        res = subprocess.run([DRIVER], input=ctl, text=True, cwd=td, env=env,
                             capture_output=True)
        if res.returncode != 0 or not os.path.exists(ofile):
            raise RuntimeError("ref_driver failed:\n" + res.stdout[-2000:] +
                               res.stderr[-2000:])
        out = np.fromfile(ofile, dtype=np.float64).reshape(-1, 9)
        t = None
        for line in res.stdout.splitlines():
            if line.startswith("loop_seconds"):
                t = float(line.split()[1])
    if return_time:
        return out, t
    return out


def ref_gauss_points(a, b, order, npts, quadr):
    """Reference composite quadrature (gaussint.f90:267-319) -> (pts, wgts)."""
    if not available():
        raise RuntimeError("reference oracle not built (make -C oracle ref)")
    res = subprocess.run([GAUSS], input=f"{a!r} {b!r} {order} {npts} {quadr}\n",
                         text=True, capture_output=True, check=True)
    arr = np.array([[float(x) for x in line.split()]
                    for line in res.stdout.splitlines() if line.strip()])
    return arr[:, 0].copy(), arr[:, 1].copy()
