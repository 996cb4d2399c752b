"""Plane-flux integration (src/fgimic/integral.f90:50-186) on the GPU.

integrate_current evaluates J at every grid point in one batched call,
weights n.(J B) with the product quadrature weights (zero outside the
radius bound around the grid centre) and returns the net, positive and
negative flux; the reference's stdout block is produced by `report`.
"""
import numpy as np



def au2si(au):
    """au -> nA/T (globals.f90:309-332)"""
    aulength = 0.52917726e-10
    auspeedoflight = 137.03599
    speedoflight = 299792458.0
    aucharge = 1.60217733e-19
    hbar = 1.05457267e-34
    autime = aulength * auspeedoflight / speedoflight
    autesla = hbar / aucharge / aulength / aulength
    audjdb = aucharge / autime / autesla
    return au * audjdb * 1.0e9


def plane_of(grid, bfield):
    return dict(points=grid.points(), weights=grid.weights(), normal=grid.normal(),
                bfield=np.asarray(bfield, dtype=np.float64), center=grid.center(),
                radius=float(grid.radius))


def integrate_current(ctx, grid, bfield, spin="total", precision="fp64"):
    """-> (net, positive, negative) flux in au through `grid`"""
    out = integrate_planes(ctx, [grid], [bfield], spin=spin, precision=precision)
    return tuple(float(v) for v in out[0])


def integrate_planes(ctx, grids, bfields, spin="total", precision="fp64"):
    """many planes in one call -> (nplanes, 3); product grids are generated
    on the device (identical points and weights), file grids go as arrays"""
    if all(g.mode != "file" for g in grids):
        return ctx.flux_plane_grids(grids, bfields, spin=spin, precision=precision)
    return ctx.flux_planes([plane_of(g, b) for g, b in zip(grids, bfields)], spin=spin,
                           precision=precision)


def report(net, pos, neg, modulus=False):
    """the reference's block for one flux (integral.f90:167-183; the
    modulus' labels from integral.f90:306-322), as its msg_out prints it
    (one leading blank, trailing blanks trimmed), with the empty records
    before and after it; a list of lines"""
    au = "Induced mod current (au)   :" if modulus else "   Induced current (au)    :"
    si = "Induced mod current (nA/T) :" if modulus else "   Induced current (nA/T)  :"
    return ["", " " + "*" * 60,
            " " + au + "%13.6f" % net,
            " " + "      Positive contribution:%13.6f  (%11.6f )" % (pos, au2si(pos)),
            " " + "      Negative contribution:%13.6f  (%11.6f )" % (neg, au2si(neg)),
            "",
            " " + si + "%13.6f" % au2si(net),
            " " + "      (conversion factor)  :%13.6f" % au2si(1.0),
            " " + "*" * 60, ""]
