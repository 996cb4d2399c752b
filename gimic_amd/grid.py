"""Grids and magnetic-field selection with the reference's conventions.

Follows src/fgimic/grid.f90 and src/fgimic/magnet.f90:
  * modes 'std'/'base' (origin, ivec, jvec, lengths; spacing or grid_points,
    grid.f90:139-159), 'bond' (grid.f90:161-266) and 'file' (grid.f90:543-576)
  * point distributions 'even' (grid.f90:351-371) and 'gauss'/'lobatto'
    (grid.f90:295-349: point counts rounded up to a multiple of the order,
    composite rule from setup_gauss_data)
  * optional rotation of the basis vectors (grid.f90:697-769)
  * point order i fastest, then j, then k (grid.f90:478-495)
  * get_magnet/check_field (magnet.f90:11-79): B flipped to point against
    the grid normal unless it is the 'X' (orthogonal) axis
  * the lines the reference prints while it builds the grid (Grid.log) and
    chooses the field (get_magnet's log), in its formats
All lengths in bohr.
"""
import math

import numpy as np

from . import lib

PII = 3.141592653589793


def _norm(v):
    v = np.asarray(v, dtype=np.float64)
    return v / math.sqrt(float(np.sum(v * v)))


def _cross(a, b):
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                     a[0] * b[1] - a[1] * b[0]], dtype=np.float64)


class Grid:
    def __init__(self):
        self.mode = "std"
        self.gtype = "even"
        self.basv = np.zeros((3, 3))   # columns are the basis vectors
        self.l = np.zeros(3)
        self.origin = np.zeros(3)
        self.ortho = np.zeros(3)
        self.step = np.ones(3)
        self.npts = np.zeros(3, dtype=int)
        self.pts = [None, None, None]
        self.wgt = [None, None, None]
        self.radius = 1.0e10
        self.gauss = False
        self.xdata = None
        self.log = []                  # stdout lines of the set-up

    # -- geometry ------------------------------------------------------------
    def points(self):
        """(N,3) grid points, i fastest (gridpoint, grid.f90:478-495)"""
        if self.mode == "file":
            return self.xdata.copy()
        p1, p2, p3 = self.pts
        k, j, i = np.meshgrid(p3, p2, p1, indexing="ij")
        r = (self.origin[None, :] + i.reshape(-1, 1) * self.basv[:, 0][None, :]
             + j.reshape(-1, 1) * self.basv[:, 1][None, :]
             + k.reshape(-1, 1) * self.basv[:, 2][None, :])
        return r

    def weights(self):
        """product quadrature weights w_i w_j w_k in point order"""
        if self.mode == "file":
            return np.ones(len(self.xdata))
        w1, w2, w3 = self.wgt
        return (w3[:, None, None] * w2[None, :, None] * w1[None, None, :]).reshape(-1)

    def gridpoint(self, i, j, k):
        """1-based indices as in the reference"""
        if self.mode == "file":
            return self.xdata[i - 1].copy()
        return (self.origin + self.pts[0][i - 1] * self.basv[:, 0]
                + self.pts[1][j - 1] * self.basv[:, 1] + self.pts[2][k - 1] * self.basv[:, 2])

    def center(self):
        """grid_center, grid.f90:517-530"""
        v1 = self.gridpoint(self.npts[0], 1, 1)
        v2 = self.gridpoint(1, self.npts[1], 1)
        return (v1 + v2) * 0.5

    def normal(self):
        return self.basv[:, 2].copy()

    @property
    def size(self):
        return int(np.prod(self.npts))


def _fmt(spec, v):
    return "".join(spec % x for x in v)


def _rotate(g, angle, reference):
    """rotate, grid.f90:697-769 (right-hand rule, z then y then x)"""
    rad = np.asarray(angle, dtype=np.float64) / 180.0 * PII
    g.log.append(" INFO: Rotation is: " + _fmt("%9.5f", rad))
    rot = np.zeros((3, 3))
    x = rad[2]
    rot[0, 0] = math.cos(x)
    rot[1, 1] = math.cos(x)
    rot[2, 2] = 1.0
    rot[0, 1] = math.sin(x)
    rot[1, 0] = -math.sin(x)
    euler = np.zeros((3, 3))
    x = rad[1]
    euler[0, 0] = math.cos(x)
    euler[1, 1] = 1.0
    euler[2, 2] = math.cos(x)
    euler[0, 2] = -math.sin(x)
    euler[2, 0] = math.sin(x)
    euler = euler @ rot
    x = rad[0]
    rot = np.zeros((3, 3))
    rot[0, 0] = 1.0
    rot[1, 1] = math.cos(x)
    rot[2, 2] = math.cos(x)
    rot[1, 2] = math.sin(x)
    rot[2, 1] = -math.sin(x)
    euler = rot @ euler
    g.origin = g.origin - reference
    g.basv = euler @ g.basv
    g.origin = euler @ g.origin
    g.origin = g.origin + reference


def _setup_even(g):
    g.gauss = False
    n1 = int(round(g.l[0] / g.step[0])) + 1
    n2 = int(round(g.l[1] / g.step[1])) + 1
    if abs(g.l[2]) < np.finfo(float).tiny or abs(g.step[2]) < np.finfo(float).tiny:
        n3 = 1
    else:
        n3 = int(round(g.l[2] / g.step[2])) + 1
    g.npts = np.array([n1, n2, n3])
    for n in range(3):
        g.pts[n] = np.arange(g.npts[n], dtype=np.float64) * g.step[n]
        g.wgt[n] = np.ones(g.npts[n])


def _setup_gauss(g, kw, quadr):
    g.log.append(" INFO: Integration grid selected.")
    order = int(kw.get("gauss_order", 7))
    if "grid_points" in kw:
        npts = [int(v) for v in kw["grid_points"]]
    else:
        spc = kw["spacing"]
        npts = [0 if (abs(spc[i]) < 1e-10 or spc[i] < 0) else int(round(g.l[i] / spc[i]))
                for i in range(3)]
    fixed = False
    for i in range(3):
        if not npts[i] > 1:
            npts[i] = 0
        rem = npts[i] % order
        if rem:
            npts[i] = npts[i] - rem + order
            fixed = True
    if fixed:
        g.log.append(" INFO: Adjusted number of grid points for quadrature: " + _fmt("%5d", npts))
    g.gauss = True
    g.npts = np.zeros(3, dtype=int)
    for i in range(3):
        if npts[i] > 0:
            g.npts[i] = npts[i]
            g.pts[i], g.wgt[i] = lib.quadrature(0.0, g.l[i], order, npts[i], quadr)
        else:
            g.npts[i] = 1
            g.pts[i], g.wgt[i] = np.zeros(1), np.ones(1)


def _nint(x):
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def new_grid(kw, coords=None, magnet=None):
    """Build a grid from Grid-section keywords (only keywords the user set,
    plus '_arg' = mode).  coords: (natoms,3) bohr, 1-based atom indices in
    'bond'/'fixpoint'.  magnet: top-level 'magnet' vector if set (it
    overrides the bond grid's orthogonal axis, grid.f90:253-256)."""
    g = Grid()
    g.mode = kw.get("_arg", "std")
    coords = None if coords is None else np.asarray(coords, dtype=np.float64)
    if g.mode == "file":
        g.gtype = "file"
        g.xdata = np.loadtxt(kw["file"]).reshape(-1, 3)
        g.npts = np.array([len(g.xdata), 1, 1])
        g.log += ["   Total number of grid points  :%10d" % g.size, ""]   # grid.f90:572-575
        return g
    if g.mode in ("std", "base"):
        g.origin = np.asarray(kw["origin"], dtype=np.float64)
        g.basv[:, 0] = kw["ivec"]
        g.basv[:, 1] = kw["jvec"]
        g.l = np.asarray(kw["lengths"], dtype=np.float64)
        if "spacing" in kw:
            g.step = np.asarray(kw["spacing"], dtype=np.float64)
        else:
            n = np.asarray(kw["grid_points"], dtype=np.float64)
            with np.errstate(divide="ignore", invalid="ignore"):
                g.step = g.l / (n - 1)
        g.basv[:, 2] = _cross(g.basv[:, 0], g.basv[:, 1])
        g.ortho = _norm(g.basv[:, 2])
    elif g.mode == "bond":
        if "bond" in kw:
            a1, a2 = kw["bond"]
            g.basv[:, 0] = coords[a1 - 1]
            g.basv[:, 1] = coords[a2 - 1]
        else:
            g.basv[:, 0] = kw["coord1"]
            g.basv[:, 1] = kw["coord2"]
        if "fixpoint" in kw:
            g.origin = coords[int(kw["fixpoint"]) - 1].copy()
        else:
            g.origin = np.asarray(kw["fixcoord"], dtype=np.float64)
        g.radius = float(kw["radius"]) if "radius" in kw else 1.0e10
        l3 = float(kw.get("distance", -1.0))
        if "height" in kw:
            hgt = [-float(kw["height"][0]), float(kw["height"][1])]
            wdt = [-float(kw["width"][0]), float(kw["width"][1])]
        else:
            wdt = [float(kw["in"]), float(kw["out"])]
            hgt = [float(kw["up"]), float(kw["down"])]
        g.l = np.array([hgt[0] + hgt[1], wdt[0] + wdt[1], 0.0])
        v1 = g.basv[:, 0] - g.origin
        v2 = g.basv[:, 1] - g.origin
        ortho = _cross(v1, v2)
        if np.all(np.abs(ortho) <= 1e-12):
            raise ValueError("Basis vectors are linearly dependent, field direction undefined!")
        g.ortho = _norm(ortho)
        v3 = _norm(v2 - v1)
        v1 = -g.ortho
        v2 = _norm(_cross(v3, v1))
        oo = g.basv[:, 0] + l3 * v3
        g.origin = oo - wdt[1] * v2 - hgt[1] * v1
        g.basv = np.stack([v1, v2, v3], axis=1)
        if magnet is not None:
            g.ortho = _norm(magnet)
        g.bond_center = oo
        # grid.f90:258-274, before any rotation
        g.log += ["", " Integration grid data", " " + "-" * 48]
        for name, v in (("center ", oo), ("origin ", g.origin), ("basv1  ", v1), ("basv2  ", v2),
                        ("basv3  ", v3), ("lenghts", g.l), ("magnet ", g.ortho)):
            g.log.append(" " + name + _fmt("%12.6f", v))
        g.log.append("")
    else:
        raise ValueError(f"Unknown grid type: {g.mode}")
    g.log.append(" Grid mode = " + g.mode)
    # normalise + ortho_coordsys (grid.f90:286-300, 398-423)
    for i in range(3):
        n = math.sqrt(float(np.sum(g.basv[:, i] ** 2)))
        if n > 0.0:
            g.basv[:, i] = g.basv[:, i] / n
    if abs(float(g.basv[:, 0] @ g.basv[:, 1])) > 1e-10:
        t = _cross(g.basv[:, 0], g.basv[:, 2])
        g.basv[:, 1] = t / math.sqrt(float(np.sum(t ** 2)))
        for i in range(3):
            n = math.sqrt(float(np.sum(g.basv[:, i] ** 2)))
            if n > 0.0:
                g.basv[:, i] = g.basv[:, i] / n
    if "rotation" in kw:
        if "rotation_origin" in kw:
            ref = np.asarray(kw["rotation_origin"], dtype=np.float64)
        else:
            if "height" in kw:
                down_length, out_length = float(kw["height"][1]), float(kw["width"][1])
            else:
                # Grid.out and Grid.down are not keywords of the input
                # template (gimic.in:90-112): the reference's getkw leaves
                # its locals unassigned (getkw.f90:598-623).  0 puts the
                # rotation reference at the grid origin, which a zero
                # rotation leaves exactly in place (test/benzene/vectors)
                out_length, down_length = float(kw.get("out", 0.0)), float(kw.get("down", 0.0))
            ref = g.origin + out_length * g.basv[:, 1] + down_length * g.basv[:, 0]
        _rotate(g, kw["rotation"], ref)
    g.gtype = kw.get("type", "even")
    if g.gtype == "even":
        _setup_even(g)
    elif g.gtype in ("gauss", "lobatto"):
        _setup_gauss(g, kw, g.gtype)
    else:
        raise ValueError(f"Unknown grid type: {g.gtype}")
    g.log += ["   Number of grid points <v1,v2>:" + _fmt("%5d", g.npts),
              "   Total number of grid points  :%10d" % g.size, ""]      # grid.f90:131-137
    return g


def get_magnet(g, magnet_axis="", magnet=None, log=None):
    """magnet.f90:11-79; `log` (a list) receives check_field's lines"""
    axis = (magnet_axis or "").strip()
    if axis in ("T",):
        axis = "X"
    ortho = False
    if axis:
        d = 1.0
        if axis[0] == "-":
            d = -1.0
            axis = axis[1:]
        c = axis[0]
        if c in "ijk":
            mag = g.basv[:, "ijk".index(c)] * d
        elif c in "xyz":
            mag = np.eye(3)["xyz".index(c)] * d
        elif c == "X":
            ortho = True
            mag = g.ortho * d
        else:
            raise ValueError(f"Invalid axis specifier: {magnet_axis}")
    else:
        mag = np.asarray(magnet, dtype=np.float64)
    if np.all(np.abs(mag) <= 1e-12):
        raise ValueError("Magnetic field is zero, not wasting more CPU.")
    mag = np.asarray(mag, dtype=np.float64)
    if not ortho:
        # a file grid has no basis vectors (grid.f90:548): x = 0, no flip
        x = float(g.basv[:, 2] @ mag)
        if x > 0.0:
            mag = -mag
        if log is not None:
            # check_field, magnet.f90:64-86 (its warning fires only for |x| > 1)
            if x > 0.0:
                log.append(" INFO: Left handed coordinate system, reversing magnetic field")
            if abs(x) - 1.0 > 1e-12 and abs(x) > 1e-12:
                log.append(" WARNING: Magnetic field not orthogonal to grid")
            log += ["    Magnetic field <x,y,z> =" + _fmt("%10.5f", mag), ""]
    return mag
