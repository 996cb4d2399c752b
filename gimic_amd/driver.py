"""fgimic-compatible driver for the current-density workloads.

Reads a gimic.inp (getkw syntax), sets up the basis and density from the
MOL/XDENS files it names, builds the grid and field (grid.f90, magnet.f90)
and runs
  calc=integral  signed-modulus flux when Essential.jmod is on, then the
                 current flux, per spin case (total; UHF adds alpha, beta,
                 spindens), then ACID when Essential.acid is on
                 -- src/fgimic/gimic.F90:222-261, integral.f90
  calc=cdens     J B on every grid point (jfield.f90:62-184)
printing the reference's stdout blocks.  All integrals come from one tensor
evaluation per spin case (the reference evaluates J once per integral).

    python -m gimic_amd.driver gimic.inp [--precision fp64|mixed]
"""
import argparse
import os
import sys

import numpy as np

from . import flux, grid as gridmod, inp
from .lib import Context


def setup(kw, basedir="."):
    kw = inp.with_defaults(kw)
    adv = kw["Advanced"]
    thrs = float(adv["screening_thrs"]) if adv.get("screening", False) else -1.0
    uhf = bool(kw.get("openshell", False))
    path = lambda f: f if os.path.isabs(f) else os.path.join(basedir, f)
    ctx = Context.from_files(path(kw["basis"]), path(kw["xdens"]), uhf=uhf, screening=thrs)
    ctx.set_flags(adv["GIAO"], adv["diamag"], adv["paramag"])
    coords, _ = ctx.atoms()
    gkw = dict(kw["Grid"])
    if "file" in gkw and not os.path.isabs(gkw["file"]):
        gkw["file"] = path(gkw["file"])
    magnet = kw.get("magnet") if "magnet" in kw and not kw.get("magnet_axis") else None
    g = gridmod.new_grid(gkw, coords, magnet=magnet)
    b = gridmod.get_magnet(g, kw.get("magnet_axis", ""), kw.get("magnet"))
    return kw, ctx, g, b, uhf


def spin_cases(uhf):
    return ["total", "alpha", "beta", "spindens"] if uhf else ["total"]


def run_integral(kw, ctx, g, b, uhf, precision="fp64", out=sys.stdout):
    jmod = bool(kw["Essential"].get("jmod", False))
    plane = flux.plane_of(g, b)
    res = {}
    for spin in spin_cases(uhf):
        cur, mod, acid = ctx.flux_planes([plane], spin=spin, precision=precision, modulus=True,
                                         acid=True)
        res[spin] = dict(current=tuple(cur[0]), modulus=tuple(mod[0]),
                         acid=float(np.sqrt(acid[0])))
    if jmod:
        print("*** Integrating |J|", file=out)
        for spin in spin_cases(uhf):
            if uhf:
                print(f"*** Integrating {spin if spin != 'spindens' else 'spin'} density", file=out)
            net, pos, neg = res[spin]["modulus"]
            print(flux.report(net, pos, neg).replace("   Induced current (au)    :",
                                                      "Induced mod current (au)   :")
                  .replace("   Induced current (nA/T)  :", "Induced mod current (nA/T) :"),
                  file=out)
    print("*** Integrating current", file=out)
    for spin in spin_cases(uhf):
        if uhf:
            print(f"*** Integrating {spin if spin != 'spindens' else 'spin'} density", file=out)
        print(flux.report(*res[spin]["current"]), file=out)
    if kw["Essential"].get("acid", False):
        # integral.f90:500-510 (the reference integrates ACID for the total density)
        a = res["total"]["acid"]
        print("*" * 60, file=out)
        print("   ACID (au) sqrt(delta J^2):%13.6f" % a, file=out)
        print("   ACID (nA/T)              :%13.6f" % flux.au2si(a), file=out)
        print("*" * 60, file=out)
    return res


def run_cdens(kw, ctx, g, b, uhf, precision="fp64"):
    pts = g.points()
    out = {}
    for spin in spin_cases(uhf):
        J = ctx.jtensor(pts, spin=spin, precision=precision)
        out[spin] = np.einsum("pbm,b->pm", J.reshape(-1, 3, 3), b)
    return pts, out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("input", nargs="?", default="gimic.inp")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    a = ap.parse_args(argv)
    kw = inp.parse(open(a.input).read())
    kw, ctx, g, b, uhf = setup(kw, os.path.dirname(os.path.abspath(a.input)))
    if kw.get("calc") == "integral":
        run_integral(kw, ctx, g, b, uhf, a.precision)
    else:
        pts, vec = run_cdens(kw, ctx, g, b, uhf, a.precision)
        for spin, v in vec.items():
            np.savetxt(f"jvec_{spin}.txt", np.hstack([pts, v]), fmt="%.10e")
    ctx.close()


if __name__ == "__main__":
    main()
