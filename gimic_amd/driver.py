"""fgimic-compatible driver for the current-density workloads.

Reads a gimic.inp (getkw syntax), sets up the basis and density from the
MOL/XDENS files it names, builds the grid and field (grid.f90, magnet.f90)
and runs
  calc=integral  signed-modulus flux when Essential.jmod is on, then the
                 current flux, per spin case (total; UHF adds alpha, beta,
                 spindens), then ACID when Essential.acid is on
                 -- src/fgimic/gimic.F90:222-261, integral.f90
  calc=cdens     J B on every grid point (jfield.f90:62-184) per spin case,
                 written as the reference writes it (jvector_plots,
                 jfield.f90:276-445): jvec[tag].vti on even grids,
                 jmod[tag].vti (signed modulus, Essential.jmod) and acid.vti
                 (Essential.acid) on 3-D grids, jmod[tag].txt on Gauss/Lobatto
                 grids, jvec.vtu on file grids when grid.1.ele is present;
                 with Essential.prop, the shielding of every nucleus and the
                 magnetizability on the external numgrid grid in coord.au,
                 gridfile.grd, grid_w.grd and nelpts.info (jfield.f90:584-929)
printing the reference's stdout blocks (the grid and field set-up lines
first); mol.xyz (write_xyz) and grid.xyz (plot_grid_xyz) in both modes.
All integrals come from one tensor evaluation per spin case (the reference
evaluates J once per integral).

    python -m gimic_amd.driver gimic.inp [--precision fp64|mixed] [--dryrun]
    python -m gimic_amd.driver - < keyword-stream     (as gimic.bin is run)

The input may be a gimic.inp or the keyword stream the reference's frontend
pipes into gimic.bin (`-` reads it from standard input; MOL, XDENS and the
output files are then relative to the working directory).
"""
import argparse
import os
import sys

import numpy as np

from . import flux, grid as gridmod, inp, writers
from .lib import Context


def setup(kw, basedir="."):
    """basis, density, grid and field of a run (gimic.F90:140-165); a dry
    run reads no density"""
    kw = inp.with_defaults(kw)
    adv = kw["Advanced"]
    thrs = float(adv["screening_thrs"]) if adv.get("screening", False) else -1.0
    uhf = bool(kw.get("openshell", False))
    path = lambda f: f if os.path.isabs(f) else os.path.join(basedir, f)
    dens = None if kw.get("dryrun", False) else path(kw["xdens"])
    ctx = Context.from_files(path(kw["basis"]), dens, uhf=uhf, screening=thrs)
    ctx.set_flags(adv["GIAO"], adv["diamag"], adv["paramag"])
    coords, _ = ctx.atoms()
    gkw = dict(kw["Grid"])
    if "file" in gkw and not os.path.isabs(gkw["file"]):
        gkw["file"] = path(gkw["file"])
    magnet = kw.get("magnet") if "magnet" in kw and not kw.get("magnet_axis") else None
    g = gridmod.new_grid(gkw, coords, magnet=magnet)
    g.magnet_log = []
    b = gridmod.get_magnet(g, kw.get("magnet_axis", ""), kw.get("magnet"), log=g.magnet_log)
    return kw, ctx, g, b, uhf


def setup_report(g, uhf, out=None):
    """what gimic.F90:159-171 prints between the density and the
    calculation: the grid's set-up, the grid plot note, the field and the
    shell type (to `out`, standard output when None)"""
    out = sys.stdout if out is None else out
    for line in g.log + [" *** Grid plot in grid.xyz"] + g.magnet_log:
        print(line, file=out)
    print(" INFO: Open-shell calculation" if uhf else " INFO: Closed-shell calculation", file=out)
    print("", file=out)


def dry_run_report(calc, out=None):
    """what a dry run prints instead of calculating (gimic.F90:174-177, and
    the headers run_cdens and run_integral print before they return,
    gimic.F90:200-204, 226-230)"""
    out = sys.stdout if out is None else out
    print(" *** Dry run, not calculating ...", file=out)
    print("", file=out)
    head = "Integrating current density" if calc == "integral" else "Calculating current density"
    print(" " + head, file=out)
    print(" " + "*" * 41, file=out)


def spin_cases(uhf):
    return ["total", "alpha", "beta", "spindens"] if uhf else ["total"]


SPIN_NAME = {"total": "total", "alpha": "alpha", "beta": "beta", "spindens": "spin"}


def integral_report(res, uhf, jmod, acid, radius, magnet_log=()):
    """the printout of an integral run (run_integral, gimic.F90:222-261):
    per spin case the |J| flux (when jmod), then the current flux, each
    integration preceded by its spin note (open shell), the field check that
    its get_magnet call prints (integral.f90:84, 224) and the radius bound
    when one is set; then ACID for the total density (integrate_acid,
    integral.f90:416-510).  res: {spin: {"modulus": (net, pos, neg),
    "current": (net, pos, neg), "acid": value}}"""
    spins = spin_cases(uhf)
    bound = []
    if radius < 1.0e10:
        bound = [" " + writers.list_record("Integration bound set to radius ", float(radius))
                 .rstrip()]

    def note(spin):
        return [f" *** Integrating {SPIN_NAME[spin]} density"] if uhf else []

    lines = [" Integrating current density", " " + "*" * 41]
    if jmod:
        lines.append(" *** Integrating |J|")
        for spin in spins:
            lines += note(spin) + list(magnet_log) + bound
            lines += flux.report(*res[spin]["modulus"], modulus=True)
        lines.append("")
    else:
        lines.append(" Jmod integration skipped.")
    lines.append(" *** Integrating current")
    for spin in spins:
        lines += note(spin) + list(magnet_log) + bound
        lines += flux.report(*res[spin]["current"])
    lines.append("")
    if acid:
        a = res["total"]["acid"]
        lines += [" *** Integrating ACID density"] + note("total") + bound
        lines += ["", " " + "*" * 60,
                  " " + "   ACID (au) sqrt(delta J^2):%13.6f" % a,
                  " " + "   ACID (nA/T)              :%13.6f" % flux.au2si(a),
                  "", " " + "*" * 60, "", ""]
    return lines


def run_integral(kw, ctx, g, b, uhf, precision="fp64", out=None):
    out = sys.stdout if out is None else out
    jmod = bool(kw["Essential"].get("jmod", False))
    # product grids are generated on the device from their 1-D rules (the
    # same points and weights, bit for bit); a file grid goes as host arrays
    plane = flux.plane_of(g, b) if g.mode == "file" else None
    res = {}
    for spin in spin_cases(uhf):
        if plane is None:
            cur, mod, acid = ctx.flux_plane_grids([g], [b], spin=spin, precision=precision,
                                                  modulus=True, acid=True)
        else:
            cur, mod, acid = ctx.flux_planes([plane], spin=spin, precision=precision,
                                             modulus=True, acid=True)
        res[spin] = dict(current=tuple(cur[0]), modulus=tuple(mod[0]),
                         acid=float(np.sqrt(acid[0])))
    for line in integral_report(res, uhf, jmod, bool(kw["Essential"].get("acid", False)),
                                g.radius, getattr(g, "magnet_log", [])):
        print(line, file=out)
    return res


def field_lines(b):
    """what jvector_plots prints before it writes a spin case's files
    (jfield.f90:303-306): the field as a list-directed record, then an
    empty string written the same way"""
    return [" magnetic field", writers.list_record(*[float(x) for x in b]).rstrip("\n"),
            writers.list_record("").rstrip("\n")]


def run_cdens(kw, ctx, g, b, uhf, precision="fp64", outdir=None, out=None):
    """J at every grid point per spin case (total; UHF adds alpha, beta and
    spindens, gimic.F90:196-220); with outdir, each case's files are written
    there as jvector_plots writes them (tags '', 'alpha', 'beta',
    'spindens'), with the reference's printout (to `out`, standard output
    when None; the reference's single-core time estimate, jfield.f90:188-229,
    is left out).  Returns {spin: (points, 9) tensors}."""
    out = sys.stdout if out is None else out
    pts = g.points()
    ess = kw["Essential"]
    elements = None
    if outdir is not None:
        print(" Calculating current density", file=out)
        print(" " + "*" * 41, file=out)
        if g.mode == "file":
            ele = os.path.join(outdir, "grid.1.ele")
            if os.path.exists(ele):
                elements = writers.read_elements(ele)
    res = {}
    for spin in spin_cases(uhf):
        res[spin] = ctx.jtensor(pts, spin=spin, precision=precision)
        if outdir is not None:
            for line in field_lines(b):
                print(line, file=out)
            if g.mode == "file" and elements is None:
                # jfield.f90:438, list-directed
                print(" not writing a vtu file, because the file grid.1.ele was not found.",
                      file=out)
            writers.jvector_plots(g, res[spin], b, tag="" if spin == "total" else spin,
                                  jmod=bool(ess.get("jmod", False)),
                                  acid=bool(ess.get("acid", False)), outdir=outdir,
                                  elements=elements)
    return res


def _fortran_a(text, w):
    """Fortran Aw output: right-justified in w columns, truncated on the right"""
    return text.rjust(w) if len(text) <= w else text[:w]


def read_property_files(basedir="."):
    """numgrid inputs of get_property (jfield.f90:612-657): nucleus
    coordinates (coord.au), grid points (gridfile.grd) and weights
    (grid_w.grd), and the points per atom block (second column of
    nelpts.info); None when any of the first three is missing (the reference
    then skips the calculation)"""
    path = lambda f: os.path.join(basedir, f)
    if not all(os.path.exists(path(f)) for f in ("coord.au", "gridfile.grd", "grid_w.grd")):
        return None
    coords = np.loadtxt(path("coord.au"), ndmin=2)[:, :3]
    pts = np.loadtxt(path("gridfile.grd"), ndmin=2)[:, :3]
    w = np.loadtxt(path("grid_w.grd"), ndmin=1).reshape(-1)
    if os.path.exists(path("nelpts.info")):
        sizes = np.loadtxt(path("nelpts.info"), ndmin=2)[:, 1].astype(np.int64)
    else:
        sizes = np.array([len(w)], dtype=np.int64)
    if len(w) != len(pts) or sizes.sum() != len(w):
        raise ValueError("gridfile.grd, grid_w.grd and nelpts.info disagree on the point count")
    return coords, pts, w, sizes


def run_property(ctx, basedir=".", spin="total", precision="fp64", out=None):
    """get_property (jfield.f90:584-929): shielding of every nucleus (ppm)
    and the magnetizability (au and SI) with positive/negative parts and
    per-atom-block contributions, in the reference's print layout (to `out`,
    standard output when None)"""
    out = sys.stdout if out is None else out
    files = read_property_files(basedir)
    if files is None:
        print(" at least one of the files coord.au, gridfile.grd, and grid_w.grd is missing."
              "Therefore any property calculation is skipped.", file=out)
        return None
    coords, pts, w, sizes = files
    res = ctx.properties(pts, w, sizes, coords, spin=spin, precision=precision)
    print(f" npts {len(w):12d}", file=out)

    def contributions(block):
        # per-block isotropic sums (the reference prints differences of
        # running totals at the block ends, which are these)
        print(" atom contributions, total, positive, negative", file=out)
        csum = np.zeros(3)
        for l, row in enumerate(block, start=1):
            v = np.array([row[:3].sum(), row[3], row[4]]) / 3.0
            print("atom %5d%14.6f%14.6f%14.6f" % (l, *v), file=out)
            csum += v
        print(_fortran_a("sum ", 10) + "%14.6f%14.6f%14.6f" % tuple(csum), file=out)
        print(" " + "*" * 52, file=out)

    for k in range(len(coords)):
        tot = res["shield"][k].sum(axis=0)
        print(f" atom {k + 1:12d}", file=out)
        print(" in ppm", file=out)
        for b, name in enumerate(("sigma_xx ", "sigma_yy ", "sigma_zz ")):
            print(" " + _fortran_a(name, 10) + "  %14.6f" % tot[b], file=out)
        print(_fortran_a("shielding constant    = ", 30) + "  %14.6f" % (tot[:3].sum() / 3.0), file=out)
        print(_fortran_a("positive contribution = ", 30) + "  %14.6f" % (tot[3] / 3.0), file=out)
        print(_fortran_a("negative contribution = ", 30) + "  %14.6f" % (tot[4] / 3.0), file=out)
        print(_fortran_a("sum = ", 30) + "  %14.6f" % ((tot[3] + tot[4]) / 3.0), file=out)
        print("  ", file=out)
        contributions(res["shield"][k])
    print("", file=out)
    tot = res["magnet"].sum(axis=0)
    print("", file=out)
    for b, name in enumerate(("chi_xx ", "chi_yy ", "chi_zz ")):
        print(" " + _fortran_a(name, 7) + "  %14.8f" % tot[b], file=out)
    iso = tot[:3].sum() / 3.0
    print(" in au", file=out)
    print(" " + _fortran_a("isotropic magnetizability chi = ", 30) + "  %14.6f" % iso, file=out)
    print(" " + _fortran_a("positive contribution         = ", 30) + "  %14.6f" % (tot[3] / 3.0), file=out)
    print(" " + _fortran_a("negative contribution         = ", 30) + "  %14.6f" % (tot[4] / 3.0), file=out)
    print(" " + _fortran_a("sum ", 30) + "  %14.6f" % ((tot[3] + tot[4]) / 3.0), file=out)
    print("", file=out)
    si = 7.89104e-29  # J/T^2 per au (jfield.f90:606)
    print(" in SI units J/T^2 ", file=out)
    print(" conversion factor: 7.89104*10^-29 J/T^2 ", file=out)
    print("", file=out)
    for name, v in (("isotropic magnetizability = ", iso), ("positive contribution     = ", tot[3] / 3.0),
                    ("negative contribution     = ", tot[4] / 3.0), ("sum ", (tot[3] + tot[4]) / 3.0)):
        print(_fortran_a(name, 30) + "  %14.6E" % (v * si), file=out)
    print(" " + "*" * 52, file=out)
    print("", file=out)
    contributions(res["magnet"])
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("input", nargs="?", default="gimic.inp",
                    help="gimic.inp or keyword stream; - reads the stream from standard input")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--dryrun", action="store_true",
                    help="set up the basis, grid and field without calculating (the frontend's -y)")
    a = ap.parse_args(argv)
    if a.input == "-":
        text, basedir = sys.stdin.read(), os.getcwd()
    else:
        text, basedir = open(a.input).read(), os.path.dirname(os.path.abspath(a.input))
    kw = inp.parse_any(text)
    if a.dryrun:
        kw["dryrun"] = True
    kw, ctx, g, b, uhf = setup(kw, basedir)
    coords, _ = ctx.atoms()
    writers.write_mol_xyz("mol.xyz", ctx.symbols(), coords)
    writers.write_grid_xyz("grid.xyz", g, ctx.symbols(), coords)
    setup_report(g, uhf)
    if kw.get("dryrun", False):
        dry_run_report(kw.get("calc"))
    elif kw.get("calc") == "integral":
        run_integral(kw, ctx, g, b, uhf, a.precision)
    else:
        run_cdens(kw, ctx, g, b, uhf, a.precision, outdir=".")
        if kw["Essential"].get("prop", False):
            run_property(ctx, basedir, precision=a.precision)
    ctx.close()


if __name__ == "__main__":
    main()
