"""The benzene tests and example of the reference, as far as they go without
a density (their XDENS is not in the tree, so no current is computed here):
everything these files print that the grid, the field and the writers alone
decide.

* The grid and field set-up lines of every benzene stdout (bond grids with
  Gauss, Lobatto and even points, spacing, radius, rotation with and
  without rotation_origin, the magnet keyword; base grids in 2-D and 3-D)
  come back byte for byte (Grid.log, get_magnet's log).
* The printout of the 12 integral runs (|J| on or skipped, ACID, the radius
  bound, the field check before every integration), rebuilt by
  driver.integral_report from the values it prints, comes back byte for
  byte around the numbers; of the two field runs, the header and the
  field block (driver.field_lines).
* The ImageData headers of every benzene VTI file (extent, origin and
  spacing printed to 17 digits) come back byte for byte.
* int-cdens/reference/jmod.txt, re-printed from its own moduli on our Gauss
  points, comes back byte for byte: 1296 points at f11.7 in angstrom pin the
  bond plane's quadrature points and the single-precision angstrom factor.
* The signed modulus of the reference's own jvec.vti equals its jmod.vti
  (to the rounding of six printed digits) with the field we choose, and the
  2d/3d runs with magnet_axis=X print the negated vectors of the runs with
  magnet=[0,0,-1]: our field for X is +z.
* examples/benzene/{int,3D}: mol.xyz and grid.xyz as written there, up to
  two conventions of the compiler that wrote them (a blank in an empty
  list-directed record, no sign on a zero): the test references, written
  by gfortran, have empty records (jmod.txt) and signed zeros (stdout).
  keyword-rotation_origin's stdout is the one test output from that other
  compiler and is compared the same way.
"""
import re

import numpy as np
import pytest

from conftest import golden, printout_mismatches, stdout_integral_lines, stdout_setup_lines
from gimic_amd import driver, grid as gridmod, inp, writers
from oracle import molio
from test_cpu_grid_flux import stdout_blocks

D = None


def fixture():
    global D
    if D is None:
        D = golden("benzene.npz")
    return D


@pytest.fixture(scope="module")
def mol(tmp_path_factory):
    p = tmp_path_factory.mktemp("benzene") / "MOL"
    p.write_text(str(fixture()["mol"]))
    return molio.read_mol(str(p))


def build(mol, text):
    kw = inp.with_defaults(inp.parse(text))
    magnet = kw.get("magnet") if "magnet" in kw and not kw.get("magnet_axis") else None
    g = gridmod.new_grid(dict(kw["Grid"]), np.asarray(mol.coords), magnet=magnet)
    mlog = []
    b = gridmod.get_magnet(g, kw.get("magnet_axis", ""), kw.get("magnet"), log=mlog)
    return kw, g, b, mlog


STDOUT_TESTS = sorted({k.split("/")[0] for k in golden("benzene.npz").files
                       if k.endswith("/stdout")})
INTEGRAL_TESTS = [n for n in STDOUT_TESTS
                  if "Integrating current density" in str(golden("benzene.npz")[f"{n}/stdout"])]
VTI_HEADS = sorted(k for k in golden("benzene.npz").files if k.endswith(".vti:head"))


def test_fixture_holds_every_benzene_output():
    assert len(STDOUT_TESTS) == 14 and len(VTI_HEADS) == 9


@pytest.mark.parametrize("name", STDOUT_TESTS)
def test_setup_lines_match_benzene_stdout(mol, name):
    """gimic.F90:159-165: new_grid's lines, the grid plot note and
    check_field's lines, between the density set-up and the shell type"""
    kw, g, b, mlog = build(mol, str(fixture()[f"{name}/input"]))
    want = stdout_setup_lines(str(fixture()[f"{name}/stdout"]))
    ours = g.log + [" *** Grid plot in grid.xyz"] + mlog
    if " " in want:
        # keyword-rotation_origin's stdout comes from the other compiler
        # (blank empty records, unsigned zeros)
        assert name == "keyword-rotation_origin"
        assert normalise_compiler(ours) == normalise_compiler(want)
    else:
        assert ours == want


def printed_results(text, uhf):
    """the flux values and ACID a printout holds, as integral_report takes them"""
    cur = stdout_blocks(text, "Induced current (au)")
    mod = stdout_blocks(text, "Induced mod current (au)")
    acid = re.findall(r"ACID \(au\) sqrt\(delta J\^2\): *([-0-9.]+)", text)
    res = {}
    for k, spin in enumerate(driver.spin_cases(uhf)):
        res[spin] = dict(current=cur[k], modulus=mod[k] if mod else None,
                         acid=float(acid[0]) if acid else None)
    return res


@pytest.mark.parametrize("name", INTEGRAL_TESTS)
def test_integral_printout_matches_benzene_stdout(mol, name):
    """run_integral's printout (gimic.F90:222-261, integral.f90) rebuilt from
    the values it prints: the layout byte for byte (|J| on or skipped, ACID,
    the radius bound, the field check of every integration) and the numbers
    to the printed digits (the nA/T figures are recomputed from six-digit au
    values: 28.2 x 5e-7)"""
    text = str(fixture()[f"{name}/stdout"])
    kw, g, b, mlog = build(mol, str(fixture()[f"{name}/input"]))
    res = printed_results(text, False)
    jmod = bool(kw["Essential"]["jmod"])
    if not jmod and "Integrating |J|" in text:
        # printed by a version that integrated |J| whatever the input said
        # (the test scripts check only the current blocks)
        jmod = True
    ours = driver.integral_report(res, False, jmod, bool(kw["Essential"]["acid"]), g.radius, mlog)
    want = stdout_integral_lines(text)
    if " " in want:
        assert name == "keyword-rotation_origin"
        ours, want = normalise_compiler(ours), normalise_compiler(want)
    assert printout_mismatches(ours, want, 1.5e-5) == []


def test_integral_printout_covers_the_cases():
    texts = {n: str(fixture()[f"{n}/stdout"]) for n in STDOUT_TESTS}
    assert "Jmod integration skipped." in texts["skip-jmod-integration"]
    assert "Integration bound set to radius    2.0000000000000000" in texts["keyword-radius"]
    assert "ACID (au) sqrt(delta J^2)" in texts["keyword-magnet"]
    assert len(INTEGRAL_TESTS) == 12


@pytest.mark.parametrize("name", [n for n in STDOUT_TESTS if n not in INTEGRAL_TESTS])
def test_field_run_printout_matches_benzene_stdout(mol, name):
    """a field run's header and jvector_plots' field block (gimic.F90:200-201,
    jfield.f90:303-306): the list-directed field vector and the blank record
    of an empty string, byte for byte (the reference's single-core time
    estimate between them is not printed here)"""
    kw, g, b, mlog = build(mol, str(fixture()[f"{name}/input"]))
    lines = str(fixture()[f"{name}/stdout"]).splitlines()
    i = lines.index(" Calculating current density")
    assert lines[i + 1] == " " + "*" * 41
    j = lines.index(" magnetic field")
    assert driver.field_lines(b) == lines[j:j + 3]
    assert lines[j + 2] == " "


def test_setup_lines_cover_the_keywords():
    """the cases above include a rotation, an adjusted point count from a
    spacing, an even bond grid (unit steps: grid_points=[40, 40, 0] gives
    11 x 8 points on the 10 x 7.2 bohr plane) and a field from the magnet
    keyword"""
    joined = {n: str(fixture()[f"{n}/stdout"]) for n in STDOUT_TESTS}
    assert "Rotation is:   0.00000  0.00000  0.52360" in joined["keyword-rotation"]
    assert "quadrature:   108   72    0" in joined["keyword-spacing"]
    assert "<v1,v2>:   11    8    1" in joined["int-grid-bond-even"]
    assert "Magnetic field <x,y,z> =   0.00000   0.00000  -1.00000" in joined["keyword-magnet"]


def test_rotation_origin_moves_the_plane(mol):
    """keyword-rotation and keyword-rotation_origin print the same set-up
    lines (the printout precedes the rotation); the planes differ by the
    rotation reference: about the plane's own corner, or about (0,0,0)"""
    _, g1, _, _ = build(mol, str(fixture()["keyword-rotation/input"]))
    _, g2, _, _ = build(mol, str(fixture()["keyword-rotation_origin/input"]))
    assert np.array_equal(g1.basv, g2.basv)
    assert not np.allclose(g1.origin, g2.origin)
    # rotation about z by 30 degrees; basv3 along -y before it
    c, s = np.cos(np.pi / 6), np.sin(np.pi / 6)
    assert np.allclose(g1.basv[:, 2], [-s, -c, 0.0], atol=1e-12)


@pytest.mark.parametrize("key", VTI_HEADS)
def test_vti_headers_match_benzene_files(mol, key, tmp_path):
    name, f = key.split("/")
    f = f.split(":")[0]
    kw, g, b, _ = build(mol, str(fixture()[f"{name}/input"]))
    path = str(tmp_path / f)
    if f == "jvec.vti":
        writers.write_vtk_vector_imagedata(path, g, np.zeros((g.size, 3)))
    else:
        writers.write_vtk_imagedata(path, g, np.zeros(g.size))
    head = str(fixture()[key])
    assert open(path).read()[:len(head)] == head


def test_jmod_txt_points_match_int_cdens(mol, tmp_path):
    kw, g, b, _ = build(mol, str(fixture()["int-cdens/input"]))
    assert writers.field_files(g) == ["jmod.txt"]
    ref = str(fixture()["int-cdens/jmod.txt"])
    rows = np.array([[float(x) for x in l.split()] for l in ref.splitlines() if l.strip()])
    assert len(rows) == g.size == 1296
    vec = np.zeros((g.size, 3))
    vec[:, 0] = rows[:, 3]
    writers.write_jmod_rows(str(tmp_path / "jmod.txt"), g.points(), vec, int(g.npts[0]))
    assert (tmp_path / "jmod.txt").read_text() == ref


def test_jmod_txt_needs_the_single_precision_factor(mol, tmp_path, monkeypatch):
    """the double 0.52917726 moves printed coordinates of this plane"""
    kw, g, b, _ = build(mol, str(fixture()["int-cdens/input"]))
    monkeypatch.setattr(writers, "AU2A", 0.52917726)
    writers.write_jmod_rows(str(tmp_path / "jmod.txt"), g.points(), np.zeros((g.size, 3)),
                            int(g.npts[0]))
    ours = [l[:33] for l in (tmp_path / "jmod.txt").read_text().splitlines()]
    ref = [l[:33] for l in str(fixture()["int-cdens/jmod.txt"]).splitlines()]
    assert ours != ref


def sample(key, ncomp):
    v = np.array([float(x) for x in str(fixture()[key]).split()])
    return v.reshape(-1, ncomp) if ncomp > 1 else v


@pytest.mark.parametrize("name,bz", [("3d", 1.0), ("3d-keyword-magnet", -1.0)])
def test_signed_modulus_of_reference_vectors(mol, name, bz):
    kw, g, b, _ = build(mol, str(fixture()[f"{name}/input"]))
    assert np.array_equal(b, [0.0, 0.0, bz])
    stride = int(fixture()["sample_stride"])
    pts = g.points()[::stride]
    v = sample(f"{name}/jvec.vti:sample", 3)
    jm = sample(f"{name}/jmod.vti:sample", 1)
    assert len(v) == len(jm) == len(pts)
    ours = writers.signed_modulus(pts, v, b)
    # six printed digits: each component and the modulus carry half a unit of
    # the sixth digit, |dv| <= sqrt(3)/2e-5 |v|_max, |djm| <= 0.5e-5 |jm|
    mag = np.abs(v).max(axis=1)
    tol = (np.sqrt(3.0) + 1.0) * 0.5e-5 * np.maximum(mag, np.abs(jm)) + 1e-300
    assert np.all(np.abs(ours - jm) <= tol)
    live = np.abs(jm) > 1e-3 * np.abs(jm).max()
    assert live.sum() > 100 and np.all(np.sign(ours[live]) == np.sign(jm[live]))


@pytest.mark.parametrize("a,b_", [("2d", "2d-keyword-magnet"), ("3d", "3d-keyword-magnet")])
def test_axis_x_field_is_opposite_to_magnet_keyword(mol, a, b_):
    va = sample(f"{a}/jvec.vti:sample", 3)
    vb = sample(f"{b_}/jvec.vti:sample", 3)
    assert np.array_equal(va, -vb)                # J linear in B: B_a = -B_b
    _, ga, ba, _ = build(mol, str(fixture()[f"{a}/input"]))
    _, gb, bb, mlog = build(mol, str(fixture()[f"{b_}/input"]))
    assert np.array_equal(ba, -bb) and np.array_equal(bb, [0.0, 0.0, -1.0])
    assert not mlog[0].startswith(" INFO: Left handed")  # kept as given


def normalise_compiler(lines):
    """a blank line as an empty record, a negative zero of F format unsigned"""
    return [re.sub(r"-(0\.0+)(?![0-9])", r" \1", l) if l.strip() else "" for l in lines]


@pytest.mark.parametrize("d", ["int", "3D"])
def test_example_mol_and_grid_xyz(mol, d, tmp_path):
    kw, g, b, _ = build(mol, str(fixture()[f"examples/{d}/gimic.inp"]))
    assert mol.symbols[0] == " C" and mol.symbols[-1] == " H"
    writers.write_mol_xyz(str(tmp_path / "mol.xyz"), mol.symbols, mol.coords)
    writers.write_grid_xyz(str(tmp_path / "grid.xyz"), g, mol.symbols, mol.coords)
    for f in ("mol.xyz", "grid.xyz"):
        ours = (tmp_path / f).read_text().splitlines()
        ref = str(fixture()[f"examples/{d}/{f}"]).splitlines()
        assert ours[1] == "" and ref[1] == " "
        assert normalise_compiler(ours) == normalise_compiler(ref)
    # the atoms on the y axis have x = -0.0 in MOL; gfortran signs the zero
    assert " C   -0.0000000000   -1.3913885041    0.0000000000" in ours
