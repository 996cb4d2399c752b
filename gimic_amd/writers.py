"""Output files of the reference's field runs, written with its layouts.

  jvec[tag].vti   J B on an even grid, PointData vectors (3e14.6) plus the
                  CellData norm of the averaged corner vectors
                  (write_vtk_vector_imagedata, src/fgimic/vtkplot.f90:88-234)
  jmod[tag].vti   signed modulus of J B (jmod2_vtkplot, jfield.f90:446-489,
                  written by write_vtk_imagedata, vtkplot.f90:14-86, e14.6)
  acid.vti        ACID Delta T^2 per point (acid_vtkplot, jfield.f90:556-582)
  jvec.vtu        J B on a file grid with the tetrahedra of grid.1.ele
                  (write_vtk_vector_unstructuredgrid, vtkplot.f90:241-316,
                  3e20.10)
  jmod[tag].txt   x, y, z (angstrom) and |J B| per point of a Gauss/Lobatto
                  grid, one block per grid row (write_jmod, jfield.f90:531-541,
                  6f11.7)
  grid.xyz        atoms and grid corners in angstrom (plot_grid_xyz,
                  grid.f90:588-670)
  mol.xyz         the atoms in angstrom (write_xyz, basis.f90:487-511)

The text layouts follow what gfortran writes for the reference's format
statements: Ew.d as 0.dddddd mantissas, list-directed records opening with a
blank, integers in 12 columns and reals in 26 (17 significant digits), a
blank between a numeric item and the next item.  `jvector_plots`
(jfield.f90:276-445) decides which files a field run writes; `field_files`
restates that choice.
"""
import math
import os

import numpy as np

# angstrom per bohr (globals.f90:51).  The reference's literal 0.52917726 has
# no D exponent, so it is a default (single precision) real that the
# REAL(DP) parameter widens: 0.52917724847793579..., which is what its
# xyz files and jmod.txt coordinates carry in their last printed digits
AU2A = float(np.float32(0.52917726))
DP33 = 0.3333333   # globals.f90:62 (0.3333333D0, a double)


# ---- Fortran edit descriptors ------------------------------------------------

def fortran_e(x, w=14, d=6):
    """Ew.d as gfortran writes it: [-]0.ddd...E+ee (+eee without the E when the
    exponent needs three digits), right-justified in w columns"""
    x = float(x)
    if x == 0.0:
        mant, e = "0" * d, 0
    else:
        s = f"{abs(x):.{d - 1}e}"
        m, ex = s.split("e")
        mant, e = m.replace(".", ""), int(ex) + 1
    if x == 0.0:
        exp = "E+00"
    elif abs(e) <= 99:
        exp = "E%s%02d" % ("-" if e < 0 else "+", abs(e))
    else:
        exp = "%s%03d" % ("-" if e < 0 else "+", abs(e))
    sign = "-" if (x < 0.0 or (x == 0.0 and math.copysign(1.0, x) < 0)) else ""
    out = f"{sign}0.{mant}{exp}"
    return out.rjust(w) if len(out) <= w else "*" * w


def fortran_f(x, w, d):
    """Fw.d (gfortran keeps the sign of values that round to zero)"""
    out = f"{float(x):{w}.{d}f}"
    return out if len(out) <= w else "*" * w


def _list_real(x):
    """one REAL(8) item of list-directed output (without its separator): 17
    significant digits in F form for 0.1 <= |x| < 1e17 (and zero), E form
    otherwise, in a 25-column field"""
    x = float(x)
    ax = abs(x)
    if x == 0.0 or 0.1 <= ax < 1e17:
        if x == 0.0:
            dec = 16
        else:
            k = math.floor(math.log10(ax)) + 1
            dec = 17 - k
            # rounding may carry into one more integer digit
            if abs(float(f"{ax:.{max(dec, 0)}f}")) >= 10.0 ** k:
                dec -= 1
        return f"{x:.{max(dec, 0)}f}".rjust(20) + " " * 5
    m, ex = f"{x:.16e}".split("e")
    return (m + "E%s%03d" % ("-" if int(ex) < 0 else "+", abs(int(ex)))).rjust(25)


def list_record(*items):
    """a list-directed WRITE(fd, *) record: a leading blank (which also
    separates the first item), then the items, character items as they are,
    a separator blank before every later numeric item and before a
    character item that follows a numeric one; integers in 11 columns"""
    out = [" "]
    prev_num = None  # None: no item yet
    for it in items:
        if isinstance(it, str):
            out.append((" " if prev_num else "") + it)
            prev_num = False
        else:
            sep = "" if prev_num is None else " "
            if isinstance(it, (int, np.integer)):
                out.append(sep + "%11d" % int(it))
            else:
                out.append(sep + _list_real(it))
            prev_num = True
    return "".join(out) + "\n"


# ---- grid geometry of the ImageData header ---------------------------------

def _image_geometry(g):
    n = [int(v) for v in g.npts]
    qmin = g.gridpoint(1, 1, 1)
    qmax = g.gridpoint(n[0], n[1], n[2])
    step = qmax - qmin
    for i in range(3):
        # vtkplot.f90:33-38: divided only where the extent is positive
        if step[i] > 1e-8:
            step[i] = step[i] / (n[i] - 1)
    return n, qmin, step


def _image_head(f, g, array_line):
    n, qmin, step = _image_geometry(g)
    f.write('<?xml version="1.0"?>\n')
    f.write(list_record('<VTKFile type="ImageData" ', 'version="0.1" byte_order="LittleEndian">'))
    f.write(list_record('  <ImageData WholeExtent="', 0, n[0] - 1, 0, n[1] - 1, 0, n[2] - 1,
                        '" Origin="', qmin[0], qmin[1], qmin[2], '" Spacing="',
                        step[0], step[1], step[2], '">'))
    f.write(list_record('  <Piece Extent="', 0, n[0] - 1, 0, n[1] - 1, 0, n[2] - 1, '">'))
    f.write(list_record('  <PointData Scalars="scalars">'))
    f.write(list_record(*array_line))
    return n


def _image_tail(f):
    for t in ('   </Piece>', '   </ImageData>', '</VTKFile>'):
        f.write(list_record(t))


def write_vtk_imagedata(fname, g, values):
    """scalar field on grid g (values in point order, i fastest), e14.6,
    four per line after a first line that holds one value"""
    v = np.asarray(values, dtype=np.float64).reshape(-1)
    with open(fname, "w") as f:
        _image_head(f, g, ('  <DataArray Name="scalars" type="Float64" ',
                           'NumberOfComponents="1" Format="ascii">'))
        # vtkplot.f90:60-70: a record ends after value 0, 4, 8, ...
        line = []
        for l, x in enumerate(v):
            line.append(fortran_e(x))
            if l % 4 == 0:
                f.write("".join(line) + "\n")
                line = []
        f.write("".join(line) + "\n")
        for t in ('   </DataArray>', '   </PointData>'):
            f.write(list_record(t))
        _image_tail(f)


def cell_norms(vec, n):
    """vtkplot.f90:166-221: per cell the norm of the mean of its corner
    vectors (8 corners, or 4 on a grid one point thick), summed in the
    reference's term order; (points, 3) in point order -> one value per
    cell, i fastest"""
    p = np.asarray(vec, dtype=np.float64).reshape(n[2], n[1], n[0], 3)
    if n[0] > 1 and n[1] > 1 and n[2] > 1:
        c = [p[:-1, :-1, :-1], p[:-1, :-1, 1:], p[:-1, 1:, :-1], p[1:, :-1, :-1],
             p[:-1, 1:, 1:], p[1:, :-1, 1:], p[1:, 1:, :-1], p[1:, 1:, 1:]]
        div = 8.0
    else:
        if n[2] == 1:
            q = p[0]
        elif n[1] == 1:
            q = p[:, 0]
        elif n[0] == 1:
            q = p[:, :, 0]
        else:
            return np.zeros(0)
        c = [q[:-1, :-1], q[:-1, 1:], q[1:, :-1], q[1:, 1:]]
        div = 4.0
    s = c[0].copy()
    for t in c[1:]:
        s = s + t
    s = s / div
    return np.sqrt(s[..., 0] ** 2 + s[..., 1] ** 2 + s[..., 2] ** 2).reshape(-1)


def write_vtk_vector_imagedata(fname, g, vec):
    """vector field (points, 3) on grid g: PointData 3e14.6 per point, CellData
    the norm of each cell's mean vector"""
    vec = np.asarray(vec, dtype=np.float64).reshape(-1, 3)
    with open(fname, "w") as f:
        n = _image_head(f, g, ('  <DataArray Name="vectors" type="Float64" ',
                               'NumberOfComponents="3" Format="ascii">'))
        f.write("".join(fortran_e(a) + fortran_e(b) + fortran_e(c) + "\n" for a, b, c in vec))
        f.write("\n")
        for t in ('   </DataArray>', '   </PointData>', '   <CellData Scalars="foo">'):
            f.write(list_record(t))
        f.write("".join(fortran_e(x) + "\n" for x in cell_norms(vec, n)))
        f.write(list_record('   </CellData>'))
        _image_tail(f)


def write_vtk_vector_unstructuredgrid(fname, pts, vec, cells):
    """file-grid vectors with the tetrahedra `cells` ((ncells, 4), 1-based
    point indices as grid.1.ele holds them); coordinates as the grid file
    gives them (bohr)"""
    pts = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    vec = np.asarray(vec, dtype=np.float64).reshape(-1, 3)
    cells = np.asarray(cells, dtype=np.int64).reshape(-1, 4)
    e = lambda x: fortran_e(x, 20, 10)
    pad = " " * 8
    with open(fname, "w") as f:
        f.write('<?xml version="1.0"?>\n')
        f.write('<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n')
        f.write('  <UnstructuredGrid>\n')
        f.write('    <Piece NumberOfPoints="%10d" NumberOfCells="%10d">\n' % (len(pts), len(cells)))
        f.write('      <Points>\n')
        f.write('        <DataArray type="Float32" NumberOfComponents="3" Format="ascii">\n')
        f.write("".join(pad + e(a) + e(b) + e(c) + "\n" for a, b, c in pts))
        f.write('        </DataArray>\n')
        f.write('      </Points>\n')
        f.write('      <PointData Scalars="scalars">\n')
        f.write('        <DataArray Name="vectors" type="Float64" NumberOfComponents="3" '
                'Format="ascii">\n')
        f.write("".join(pad + e(a) + e(b) + e(c) + "\n" for a, b, c in vec))
        f.write('        </DataArray>\n')
        f.write('      </PointData>\n')
        f.write('      <Cells>\n')
        f.write('        <DataArray type="Int32" Name="connectivity" Format="ascii">\n')
        f.write("".join(pad + "%10d%10d%10d%10d\n" % tuple(c - 1) for c in cells))
        f.write('        </DataArray>\n')
        f.write('        <DataArray type="Int32" Name="offsets" Format="ascii">\n')
        f.write(pad + "".join("%10d" % (4 * (c + 1)) for c in range(len(cells))) + "\n")
        f.write('        </DataArray>\n')
        f.write('        <DataArray type="Int32" Name="types" Format="ascii">\n')
        f.write(pad + "%5d" % 10 * len(cells) + "\n")
        f.write('        </DataArray>\n')
        f.write('      </Cells>\n')
        f.write('      <CellData Scalars="foo">\n')
        f.write(pad + " 0.0" * len(cells) + "\n")
        f.write('      </CellData>\n')
        f.write('    </Piece>\n')
        f.write('  </UnstructuredGrid>\n')
        f.write('</VTKFile>\n')


def read_elements(fname):
    """grid.1.ele (TetGen): a count line, then per tetrahedron an index and
    four 1-based point indices (jfield.f90:419-428)"""
    with open(fname) as f:
        ncells = int(f.readline().split()[0])
        cells = np.loadtxt(f, dtype=np.int64, max_rows=ncells, ndmin=2)
    return cells[:, 1:5]


def write_jmod_rows(fname, pts_bohr, vec, row_len):
    """jmod.txt: x, y, z in angstrom and |v| (6f11.7, four values used) per
    point, an empty record after every `row_len` points (jfield.f90:364-377)"""
    pts = np.asarray(pts_bohr, dtype=np.float64).reshape(-1, 3) * AU2A
    mod = np.sqrt(np.sum(np.asarray(vec, dtype=np.float64).reshape(-1, 3) ** 2, axis=1))
    with open(fname, "w") as f:
        for p in range(len(pts)):
            f.write(fortran_f(pts[p, 0], 11, 7) + fortran_f(pts[p, 1], 11, 7) +
                    fortran_f(pts[p, 2], 11, 7) + fortran_f(mod[p], 11, 7) + "\n")
            if row_len and (p + 1) % row_len == 0:
                f.write("\n")


def signed_modulus(pts_bohr, vec, b):
    """jmod2_vtkplot (jfield.f90:446-489): |v| with the sign of
    (b x (r - (b.r) b)) . v"""
    r = np.asarray(pts_bohr, dtype=np.float64).reshape(-1, 3)
    v = np.asarray(vec, dtype=np.float64).reshape(-1, 3)
    b = np.asarray(b, dtype=np.float64)
    val = np.sqrt(np.sum(v * v, axis=1))
    c = r - (r @ b)[:, None] * b[None, :]
    nrm = np.cross(b[None, :], c)
    sgn = np.sum(nrm * v, axis=1)
    return np.where(sgn < 0.0, -val, val)


def acid_values(jt):
    """get_acid (acid.f90:11-40) per point of (points, 9) tensors"""
    t = np.asarray(jt, dtype=np.float64).reshape(-1, 9)
    d1 = (t[:, 0] - t[:, 4]) * (t[:, 0] - t[:, 4])
    d2 = (t[:, 4] - t[:, 8]) * (t[:, 4] - t[:, 8])
    d3 = (t[:, 8] - t[:, 0]) * (t[:, 8] - t[:, 0])
    o1 = (t[:, 3] + t[:, 1]) * (t[:, 3] + t[:, 1])
    o2 = (t[:, 6] + t[:, 2]) * (t[:, 6] + t[:, 2])
    o3 = (t[:, 7] + t[:, 5]) * (t[:, 7] + t[:, 5])
    return DP33 * (d1 + d2 + d3) + 0.5 * (o1 + o2 + o3)


def _xyz_line(symbol, r_bohr):
    """'(a, 3f16.10)' of a symbol and a position converted to angstrom"""
    return symbol + "".join(fortran_f(x * AU2A, 16, 10) for x in r_bohr) + "\n"


def write_mol_xyz(fname, symbols, coords):
    """write_xyz (basis.f90:487-511), written when the basis is read
    (basis.f90:84): the atom count as a list-directed record, an empty
    record, then each atom's two-character symbol and position in angstrom"""
    coords = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
    with open(fname, "w") as f:
        f.write(list_record(len(coords)))
        f.write("\n")
        for s, r in zip(symbols, coords):
            f.write(_xyz_line(s[:2].ljust(2), r))


def write_grid_xyz(fname, g, symbols, coords):
    """plot_grid_xyz (grid.f90:588-670): atoms, the grid's corners as 'X'
    and a 'Be' marker two bohr along the grid normal (std/base) or the
    orthogonal axis (bond), in angstrom"""
    coords = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
    p1, p2, p3 = (int(v) for v in g.npts)
    line = _xyz_line
    nat = len(coords)
    with open(fname, "w") as f:
        if p3 > 1:
            f.write(list_record(nat + 8 + 1))
        elif p3 == 1:
            f.write(list_record(nat + 4 + 1))
        else:
            f.write(list_record(nat))
        f.write("\n")
        for s, r in zip(symbols, coords):
            f.write(line(s[:2].ljust(2), r))
        if p3 > 1:
            corners = [(1, 1, 1), (p1, 1, 1), (1, p2, 1), (1, 1, p3), (p1, p2, 1), (p1, 1, p3),
                       (1, p2, p3), (p1, p2, p3)]
        elif p3 == 1:
            corners = [(1, 1, 1), (p1, 1, 1), (1, p2, 1), (p1, p2, 1)]
        else:
            corners = []
        for c in corners:
            f.write(line("X ", g.gridpoint(*c)))
        if g.mode in ("std", "base"):
            f.write(line("Be ", g.origin + g.basv[:, 2] * 2.0))
        elif g.mode == "bond":
            f.write(line("Be ", g.origin + g.ortho * 2.0))


def is_3d(g):
    """grid_is_3d, grid.f90:837-845"""
    return all(int(v) > 1 for v in g.npts)


def field_files(g, tag="", jmod=False, acid=False, have_elements=False):
    """names of the files jvector_plots (jfield.f90:276-445) writes for one
    spin case of a field run on grid g"""
    out = []
    if g.gauss:
        out.append(f"jmod{tag}.txt")
    if is_3d(g):
        if acid:
            out.append("acid.vti")
        if jmod:
            out.append(f"jmod{tag}.vti")
    if g.mode in ("std", "base", "bond") and g.gtype == "even":
        out.append(f"jvec{tag}.vti")
    elif (g.mode in ("std", "base") and g.gauss) or g.mode == "file":
        if have_elements:
            out.append("jvec.vtu")
    return out


def jvector_plots(g, jt, b, tag="", jmod=False, acid=False, outdir=".", elements=None):
    """write one spin case's field files the way jvector_plots does
    (jfield.f90:276-445): jt (points, 9) tensors in grid order, b the field;
    elements: (ncells, 4) tetrahedra of a file grid (grid.1.ele) or None.
    Returns the paths written."""
    jt = np.asarray(jt, dtype=np.float64).reshape(-1, 9)
    b = np.asarray(b, dtype=np.float64)
    # v = J b (compute_jvectors, jfield.f90:174-184): v_m = sum_b J(m, b) b_b
    vec = jt[:, 0:3] * b[0] + jt[:, 3:6] * b[1] + jt[:, 6:9] * b[2]
    pts = g.points()
    written = []
    path = lambda name: os.path.join(outdir, name)
    if g.mode == "bond" and not is_3d(g) and g.radius > 0.1:
        # jfield.f90:330-349, 361-371: points farther than the radius from
        # the grid centre are zeroed; the reference measures that distance
        # from the point in angstrom to the centre in bohr
        centre = g.center()
        r = np.sqrt(np.sum((pts * AU2A - centre[None, :]) ** 2, axis=1))
        vec = np.where((r > g.radius)[:, None], 0.0, vec)
    if g.gauss and g.mode in ("std", "base", "bond"):
        write_jmod_rows(path(f"jmod{tag}.txt"), pts, vec, int(g.npts[0]))
        written.append(path(f"jmod{tag}.txt"))
    if is_3d(g):
        if acid:
            write_vtk_imagedata(path("acid.vti"), g, acid_values(jt))
            written.append(path("acid.vti"))
        if jmod:
            write_vtk_imagedata(path(f"jmod{tag}.vti"), g, signed_modulus(pts, vec, b))
            written.append(path(f"jmod{tag}.vti"))
    if g.mode in ("std", "base", "bond") and g.gtype == "even":
        write_vtk_vector_imagedata(path(f"jvec{tag}.vti"), g, vec)
        written.append(path(f"jvec{tag}.vti"))
    elif g.mode == "file" and elements is not None:
        write_vtk_vector_unstructuredgrid(path("jvec.vtu"), pts, vec, elements)
        written.append(path("jvec.vtu"))
    return written
