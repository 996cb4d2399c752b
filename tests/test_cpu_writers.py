"""Field-file writers (gimic_amd/writers.py) against the reference's own
output files and the arithmetic the files carry.

* The reference's files, re-printed from their own values, come back byte
  for byte (the edit descriptors and list-directed records are right).
* The oracle's tensors, written by these writers, reproduce the reference's
  open-shell/3d jvec*.vti and jmod*.vti and c4h4/read-grid jvec.vtu byte for
  byte, and pass runtest's comparison (rel 1e-7 and 1e-8 as the reference's
  test scripts set them).
* Cell norms, the signed modulus and ACID against closed forms.
"""
import os

import numpy as np
import pytest

from conftest import golden, runtest_mismatches, write_fixture, writer_fixture
from gimic_amd import grid as gridmod, inp, writers
from oracle import molio, oracle

TAGS = ["", "alpha", "beta", "spindens"]
SPIN_OF = {"": "total", "alpha": "alpha", "beta": "beta", "spindens": "spindens"}


def test_edit_descriptors():
    e = writers.fortran_e
    assert e(0.0) == "  0.000000E+00"
    assert e(6.48806e-14) == "  0.648806E-13"
    assert e(-6.71910e-14) == " -0.671910E-13"
    assert e(0.99999996) == "  0.100000E+01"          # rounding carries into the exponent
    assert e(1.5e-120) == "  0.150000-119"            # three-digit exponents drop the E
    assert e(123.456789, 20, 10) == "    0.1234567890E+03"
    assert writers.fortran_f(-1e-9, 11, 7) == " -0.0000000"
    # list-directed records: a leading blank, I11 integers, 25-column reals
    assert writers.list_record(5) == "           5\n"
    assert writers.list_record("a", 0, "b") == " a           0 b\n"
    assert writers.list_record(1.0) == "   1.0000000000000000     \n"
    assert writers.list_record(0.5) == "  0.50000000000000000     \n"


@pytest.mark.parametrize("name", ["jvec", "jmod"])
def test_reprinted_reference_values_give_the_reference_file(name):
    """parse a reference .vti and write its values again: the same bytes
    (the CellData of jvec holds norms of full-precision vectors, so there
    only the point part and the tail are compared)"""
    ref = writer_fixture(f"open-shell/3d/reference/{name}.vti")
    kw = inp.with_defaults(inp.parse(str(golden("openshell.npz")["threed_inp"])))
    g = gridmod.new_grid(dict(kw["Grid"]), np.zeros((1, 3)))
    body = ref[ref.index('ascii">\n') + 8:ref.index("\n\n    </DataArray>")]
    vals = np.array([float(x) for x in body.split()])
    out_path = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"reprint_{os.getpid()}_{name}.vti")
    try:
        if name == "jvec":
            writers.write_vtk_vector_imagedata(out_path, g, vals.reshape(-1, 3))
            out = open(out_path).read()
            assert out[:out.index("<CellData")] == ref[:ref.index("<CellData")]
            assert out[out.index("</CellData>"):] == ref[ref.index("</CellData>"):]
        else:
            writers.write_vtk_imagedata(out_path, g, vals)
            assert open(out_path).read() == ref
    finally:
        if os.path.exists(out_path):
            os.remove(out_path)


@pytest.fixture(scope="module")
def openshell_oracle(tmp_path_factory):
    td = str(tmp_path_factory.mktemp("openshell_writers"))
    mol, xd = write_fixture(td, "openshell.npz")
    b = molio.read_mol(mol)
    dens = molio.read_xdens(xd, b, uhf=True)
    return td, b, dens


def test_oracle_fields_reproduce_openshell_3d_files(openshell_oracle):
    """test/open-shell/3d: 33^3 even grid, magnet_axis=X, jmod=on; four spin
    cases, eight files"""
    td, b, dens = openshell_oracle
    kw = inp.with_defaults(inp.parse(str(golden("openshell.npz")["threed_inp"])))
    g = gridmod.new_grid(dict(kw["Grid"]), b.coords)
    B = gridmod.get_magnet(g, kw["magnet_axis"], kw.get("magnet"))
    pts = g.points()
    for tag in TAGS:
        J = oracle.jtensor(b.arrays(), dens, pts, screening=1e-8, spin=SPIN_OF[tag])
        written = writers.jvector_plots(g, J, B, tag=tag, jmod=True, acid=False, outdir=td)
        assert sorted(os.path.basename(p) for p in written) == sorted(
            writers.field_files(g, tag, jmod=True, acid=False))
        for name in (f"jvec{tag}.vti", f"jmod{tag}.vti"):
            out = open(os.path.join(td, name)).read()
            ref = writer_fixture(f"open-shell/3d/reference/{name}")
            assert runtest_mismatches(out, ref, 1e-7) == [], name
            assert out == ref, name


def test_oracle_field_reproduces_readgrid_vtu(tmp_path):
    """test/c4h4/read-grid: file grid, B = (0, 0, -1), tetrahedra from
    grid.1.ele, jvec.vtu compared at rel 1e-8 and byte for byte"""
    d = golden("c4h4.npz")
    mol, xd = write_fixture(str(tmp_path), "c4h4.npz")
    b = molio.read_mol(mol)
    dens = molio.read_xdens(xd, b)
    ele = tmp_path / "grid.1.ele"
    ele.write_text(writer_fixture("c4h4/read-grid/grid.1.ele"))
    cells = writers.read_elements(str(ele))
    assert cells.shape == (23221, 4) and cells.min() == 1
    pts = d["readgrid_points"]
    J = oracle.jtensor(b.arrays(), dens, pts, screening=1e-8)
    g = gridmod.Grid()
    g.mode, g.gtype, g.xdata = "file", "file", pts
    g.npts = np.array([len(pts), 1, 1])
    written = writers.jvector_plots(g, J, np.array([0.0, 0.0, -1.0]), outdir=str(tmp_path),
                                    elements=cells)
    assert [os.path.basename(p) for p in written] == ["jvec.vtu"]
    out = (tmp_path / "jvec.vtu").read_text()
    ref = writer_fixture("c4h4/read-grid/reference/jvec.vtu")
    assert runtest_mismatches(out, ref, 1e-8) == []
    assert out == ref


def test_runtest_comparison_catches_a_changed_digit():
    ref = writer_fixture("open-shell/3d/reference/jmod.vti")
    i = ref.index("0.269130E-12")
    bad = ref[:i] + "0.269131E-12" + ref[i + 12:]
    assert runtest_mismatches(bad, ref, 1e-7) == [("0.269131E-12", "0.269130E-12")]
    with pytest.raises(AssertionError):
        runtest_mismatches(ref + " 1.0", ref, 1e-7)


def test_cell_norms_closed_forms():
    n = [4, 3, 5]
    npt = int(np.prod(n))
    const = np.tile([3.0, -4.0, 12.0], (npt, 1))
    assert np.allclose(writers.cell_norms(const, n), 13.0, rtol=0, atol=1e-14)
    assert writers.cell_norms(const, n).shape == (3 * 2 * 4,)
    # a linear field: each cell mean is the field at the cell centre
    k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    lin = np.stack([i.ravel() * 1.0, j.ravel() * 2.0, np.zeros(npt)], 1)
    kc, jc, ic = np.meshgrid(np.arange(n[2] - 1) + 0.5, np.arange(n[1] - 1) + 0.5,
                             np.arange(n[0] - 1) + 0.5, indexing="ij")
    want = np.hypot(ic.ravel(), 2.0 * jc.ravel())
    assert np.allclose(writers.cell_norms(lin, n), want, rtol=1e-15, atol=0)
    # grids one point thick average four corners; a line has no cells
    assert writers.cell_norms(np.ones((12, 3)), [4, 3, 1]).shape == (6,)
    assert writers.cell_norms(np.ones((12, 3)), [4, 1, 3]).shape == (6,)
    assert writers.cell_norms(np.ones((12, 3)), [1, 4, 3]).shape == (6,)
    assert writers.cell_norms(np.ones((4, 3)), [4, 1, 1]).shape == (0,)


def test_signed_modulus_and_acid_closed_forms():
    b = np.array([0.0, 0.0, 1.0])
    r = np.array([[1.0, 0.0, 0.3], [0.0, 2.0, -1.0]])
    # a diamagnetic ring current b x r runs along b x r: positive; reversed, negative
    v = np.cross(b, r)
    assert np.allclose(writers.signed_modulus(r, v, b), np.linalg.norm(v, axis=1))
    assert np.allclose(writers.signed_modulus(r, -v, b), -np.linalg.norm(v, axis=1))
    # ACID: zero for an isotropic and for an antisymmetric tensor, 2/3 (with
    # the reference's 0.3333333) for diag(1, 0, 0), 1/2 * 4 for xy = yx = 1
    iso = np.eye(3).reshape(1, 9) * 2.5
    anti = np.array([[0.0, 1.0, -2.0], [-1.0, 0.0, 3.0], [2.0, -3.0, 0.0]]).T.reshape(1, 9)
    d1 = np.diag([1.0, 0.0, 0.0]).reshape(1, 9)
    sym = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 0.0]]).reshape(1, 9)
    assert writers.acid_values(iso)[0] == 0.0
    assert writers.acid_values(anti)[0] == 0.0
    assert writers.acid_values(d1)[0] == 2 * 0.3333333
    assert writers.acid_values(sym)[0] == 2.0


def _bond_grid(kind, coords, radius=None):
    kw = {"_arg": "bond", "type": kind, "gauss_order": 5, "bond": [1, 2],
          "fixcoord": [0.0, 0.0, 5.0], "distance": 0.5, "height": [-2.0, 2.0],
          "width": [-1.0, 3.0], "grid_points": [10, 15, 0]}
    if kind == "even":
        kw["spacing"] = [0.5, 0.5, 0.5]
        kw.pop("grid_points")
    if radius is not None:
        kw["radius"] = radius
    return gridmod.new_grid(kw, coords)


def test_jmod_txt_rows_and_radius(tmp_path):
    coords = np.array([[0.0, 0.0, 0.0], [1.4, 0.3, 0.0], [0.2, 2.0, 0.1]])
    g = _bond_grid("lobatto", coords)
    assert writers.field_files(g) == ["jmod.txt"]
    npt = g.size
    jt = np.random.default_rng(1).standard_normal((npt, 9))
    B = np.array([0.0, 0.0, 1.0])
    written = writers.jvector_plots(g, jt, B, outdir=str(tmp_path))
    assert [os.path.basename(p) for p in written] == ["jmod.txt"]
    lines = (tmp_path / "jmod.txt").read_text().split("\n")
    rows = [l for l in lines if l.strip()]
    assert len(rows) == npt and lines.count("") == int(g.npts[1]) + 1  # a blank per row, final newline
    vals = np.array([[float(x) for x in l.split()] for l in rows])
    v = jt[:, 6:9]
    assert np.allclose(vals[:, :3], g.points() * writers.AU2A, atol=6e-8)
    assert np.allclose(vals[:, 3], np.linalg.norm(v, axis=1), atol=6e-8)
    # with a radius, points farther than it from the grid centre are zeroed,
    # the distance taken from the point in angstrom to the centre in bohr
    g2 = _bond_grid("lobatto", coords, radius=1.5)
    writers.jvector_plots(g2, jt, B, outdir=str(tmp_path))
    rows2 = [l for l in (tmp_path / "jmod.txt").read_text().split("\n") if l.strip()]
    mod2 = np.array([float(l.split()[3]) for l in rows2])
    far = np.linalg.norm(g2.points() * writers.AU2A - g2.center(), axis=1) > 1.5
    assert far.any() and (~far).any()
    assert np.all(mod2[far] == 0.0) and np.all(mod2[~far] > 0.0)


def test_files_per_grid_kind(tmp_path):
    coords = np.array([[0.0, 0.0, 0.0], [1.4, 0.3, 0.0], [0.2, 2.0, 0.1]])
    g = _bond_grid("even", coords)
    assert writers.field_files(g, "alpha", jmod=True, acid=True) == ["jvecalpha.vti"]
    cube = gridmod.new_grid({"_arg": "base", "origin": [-1.0, -1.0, -1.0], "ivec": [1, 0, 0],
                             "jvec": [0, 1, 0], "lengths": [2.0, 2.0, 2.0],
                             "spacing": [0.5, 0.5, 0.5]})
    assert writers.field_files(cube, "", jmod=True, acid=True) == ["acid.vti", "jmod.vti",
                                                                  "jvec.vti"]
    jt = np.random.default_rng(2).standard_normal((cube.size, 9))
    writers.jvector_plots(cube, jt, np.array([0.0, 0.0, 1.0]), jmod=True, acid=True,
                          outdir=str(tmp_path))
    txt = (tmp_path / "acid.vti").read_text()
    body = txt[txt.index('ascii">\n') + 8:txt.index("\n\n    </DataArray>")]
    acid = np.array([float(x) for x in body.split()])
    want = writers.acid_values(jt)
    # e14.6 keeps six significant digits: half a unit of the sixth is at most
    # 5e-6 of the value
    assert np.all(np.abs(acid - want) <= 5.0001e-6 * np.abs(want))


def test_grid_xyz(tmp_path):
    cube = gridmod.new_grid({"_arg": "base", "origin": [-1.0, -1.0, -1.0], "ivec": [1, 0, 0],
                             "jvec": [0, 1, 0], "lengths": [2.0, 2.0, 2.0],
                             "spacing": [0.5, 0.5, 0.5]})
    coords = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    writers.write_grid_xyz(str(tmp_path / "grid.xyz"), cube, ["C ", "H1"], coords)
    lines = (tmp_path / "grid.xyz").read_text().splitlines()
    assert int(lines[0]) == 2 + 8 + 1 and lines[1] == ""
    assert lines[2].startswith("C ") and lines[3].startswith("H1")
    corners = np.array([[float(x) for x in l[2:].split()] for l in lines[4:12]])
    assert np.allclose(np.abs(corners), writers.AU2A, atol=1e-10)
    assert lines[12].startswith("Be ") and len(lines) == 13
