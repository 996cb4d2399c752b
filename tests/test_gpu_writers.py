"""The fgimic-compatible driver on the GPU writes the reference's field files.

`python -m gimic_amd.driver gimic.inp` (driver.main, run in a scratch
directory) with the reference's own test inputs:
  * test/open-shell/3d (33^3 even grid, UHF, jmod=on): jvec*.vti and
    jmod*.vti of all four spin cases against the reference's files with
    runtest's comparison at rel 1e-7, as the reference's test script sets it;
    with acid=on also acid.vti, against the same file written from the
    oracle's tensors;
  * test/c4h4/read-grid (file grid, B = (0, 0, -1)): jvec.vtu at rel 1e-8;
  * grid.xyz in both runs, mol.xyz against the oracle's reading of MOL;
  * test/{c4h4,open-shell}/integration: the set-up lines and flux blocks of
    the printout.
"""

import numpy as np
import pytest

from conftest import golden, runtest_mismatches, write_fixture, writer_fixture
from gimic_amd import driver, writers
from oracle import molio, oracle

pytestmark = pytest.mark.gpu

READGRID_INP = """calc=cdens
basis="MOL"
xdens="XDENS"
openshell=false
magnet=[0.0, 0.0, -1.0]
Grid(file) {
    file=gridfile.grd
}
Advanced {
    spherical=off
    diamag=on
    paramag=on
    GIAO=on
    screening=on
    screening_thrs=1.d-8
}
"""


def _run_driver(tmp_path, monkeypatch, inp_text):
    (tmp_path / "gimic.inp").write_text(inp_text)
    monkeypatch.chdir(tmp_path)
    driver.main([str(tmp_path / "gimic.inp")])


def test_driver_writes_openshell_3d_files(tmp_path, monkeypatch, capsys):
    write_fixture(str(tmp_path), "openshell.npz")
    inp_text = str(golden("openshell.npz")["threed_inp"])
    print("case driver open-shell/3d cdens", flush=True)
    _run_driver(tmp_path, monkeypatch, inp_text)
    # the printout: the run's header, then per spin case the field block of
    # jvector_plots (jfield.f90:303-306)
    lines = capsys.readouterr().out.splitlines()
    i = lines.index(" Calculating current density")
    assert lines[i + 1] == " " + "*" * 41
    blocks = [k for k, l in enumerate(lines) if l == " magnetic field"]
    assert len(blocks) == 4
    for k in blocks:
        assert len(lines[k + 1].split()) == 3 and lines[k + 2] == " "
    for tag in ("", "alpha", "beta", "spindens"):
        for name in (f"jvec{tag}.vti", f"jmod{tag}.vti"):
            out = (tmp_path / name).read_text()
            ref = writer_fixture(f"open-shell/3d/reference/{name}")
            assert runtest_mismatches(out, ref, 1e-7) == [], name
    assert not (tmp_path / "acid.vti").exists()
    lines = (tmp_path / "grid.xyz").read_text().splitlines()
    assert int(lines[0]) == len(lines) - 2
    # mol.xyz from the library's reading of MOL (gimic_ctx_atom_symbols,
    # gimic_ctx_atoms) equals the one written from the oracle's reading
    b = molio.read_mol(str(tmp_path / "MOL"))
    writers.write_mol_xyz(str(tmp_path / "mol_oracle.xyz"), b.symbols, b.coords)
    assert (tmp_path / "mol.xyz").read_text() == (tmp_path / "mol_oracle.xyz").read_text()


def test_driver_acid_vti_matches_oracle_written_file(tmp_path, monkeypatch):
    mol, xd = write_fixture(str(tmp_path), "openshell.npz")
    inp_text = str(golden("openshell.npz")["threed_inp"]).replace("acid=off", "acid=on")
    assert "acid=on" in inp_text
    print("case driver open-shell/3d acid.vti", flush=True)
    _run_driver(tmp_path, monkeypatch, inp_text)
    # every spin case rewrites acid.vti (jfield.f90:388-392): the last,
    # spindens, is what stays
    from gimic_amd import grid as gridmod, inp
    b = molio.read_mol(mol)
    dens = molio.read_xdens(xd, b, uhf=True)
    kw = inp.with_defaults(inp.parse(inp_text))
    g = gridmod.new_grid(dict(kw["Grid"]), b.coords)
    J = oracle.jtensor(b.arrays(), dens, g.points(), screening=1e-8, spin="spindens")
    ref_dir = tmp_path / "oracle_files"
    ref_dir.mkdir()
    writers.write_vtk_imagedata(str(ref_dir / "acid.vti"), g, writers.acid_values(J))
    out = (tmp_path / "acid.vti").read_text()
    ref = (ref_dir / "acid.vti").read_text()
    assert runtest_mismatches(out, ref, 1e-7) == []


def test_driver_writes_readgrid_vtu(tmp_path, monkeypatch):
    write_fixture(str(tmp_path), "c4h4.npz")
    d = golden("c4h4.npz")
    np.savetxt(tmp_path / "gridfile.grd", d["readgrid_points"], fmt="%.17g")
    (tmp_path / "grid.1.ele").write_text(writer_fixture("c4h4/read-grid/grid.1.ele"))
    print("case driver c4h4/read-grid cdens", flush=True)
    _run_driver(tmp_path, monkeypatch, READGRID_INP)
    out = (tmp_path / "jvec.vtu").read_text()
    ref = writer_fixture("c4h4/read-grid/reference/jvec.vtu")
    assert runtest_mismatches(out, ref, 1e-8) == []
    assert (tmp_path / "grid.xyz").exists()


@pytest.mark.parametrize("name", ["c4h4.npz", "openshell.npz"])
def test_driver_integral_run_prints_reference_lines(tmp_path, monkeypatch, capsys, name):
    """test/{c4h4,open-shell}/integration through driver.main: the set-up
    lines byte for byte; open-shell's whole integral printout (jmod=on, four
    spin cases) around the numbers byte for byte, the numbers to the printed
    digits (one unit of the sixth decimal in au, 28.2 times that in nA/T);
    c4h4's input leaves jmod off (its printout is from a version that
    integrated |J| unasked): its current block"""
    from conftest import printout_mismatches, stdout_integral_lines, stdout_setup_lines
    from test_cpu_grid_flux import stdout_blocks
    write_fixture(str(tmp_path), name)
    d = golden(name)
    print(f"case driver {name} integration", flush=True)
    _run_driver(tmp_path, monkeypatch, str(d["integration_inp"]))
    out = capsys.readouterr().out
    ref = str(d["integration_stdout"])
    assert stdout_setup_lines(out) == stdout_setup_lines(ref)
    if name == "openshell.npz":
        assert printout_mismatches(stdout_integral_lines(out), stdout_integral_lines(ref),
                                   3e-5) == []
    else:
        assert " Jmod integration skipped." in out.splitlines()
        ours = stdout_blocks(out, "Induced current (au)")
        want = stdout_blocks(ref, "Induced current (au)")
        assert len(ours) == len(want) == 1
        assert np.allclose(ours, want, rtol=0, atol=1.01e-6), (ours, want)


def test_driver_dry_run_reads_no_density(tmp_path, monkeypatch, capsys):
    """dryrun=true (the frontend's -y): the basis, grid and field are set up,
    mol.xyz and grid.xyz written and the set-up lines printed as in a full
    run, then the dry-run note and the calculation's header; the density
    file is never opened (it is removed here) and nothing is calculated"""
    from conftest import stdout_setup_lines
    write_fixture(str(tmp_path), "c4h4.npz")
    (tmp_path / "XDENS").unlink()
    d = golden("c4h4.npz")
    text = str(d["integration_inp"])
    print("case driver c4h4 dry run", flush=True)
    _run_driver(tmp_path, monkeypatch, "dryrun=true\n" + text)
    out = capsys.readouterr().out
    assert stdout_setup_lines(out) == stdout_setup_lines(str(d["integration_stdout"]))
    lines = out.splitlines()
    i = lines.index(" *** Dry run, not calculating ...")
    assert lines[i + 1:i + 4] == ["", " Integrating current density", " " + "*" * 41]
    assert "Induced current" not in out
    assert (tmp_path / "mol.xyz").exists() and (tmp_path / "grid.xyz").exists()


# READGRID_INP as the reference's frontend pipes it into gimic.bin
READGRID_STREAM = """SECT top 2 True
ARG F KW 5
STR calc 1 True
cdens
STR basis 1 True
MOL
STR xdens 1 True
XDENS
BOOL openshell 1 True
False
DBL_ARRAY magnet 3 True
0.0
0.0
-1.0
SECT Grid 0 True
ARG T KW 1
STR Grid 1 True
file
STR file 1 True
gridfile.grd
SECT Advanced 0 True
ARG F KW 6
BOOL spherical 1 True
False
BOOL diamag 1 True
True
BOOL paramag 1 True
True
BOOL GIAO 1 True
True
BOOL screening 1 True
True
DBL screening_thrs 1 True
1.d-8
"""


def test_driver_reads_keyword_stream_on_stdin(tmp_path, monkeypatch):
    """`python -m gimic_amd.driver - < stream`, as gimic.bin is run: the same
    jvec.vtu as from the gimic.inp"""
    import io
    from gimic_amd import inp
    assert inp.parse_stream(READGRID_STREAM) == inp.parse(READGRID_INP)
    write_fixture(str(tmp_path), "c4h4.npz")
    d = golden("c4h4.npz")
    np.savetxt(tmp_path / "gridfile.grd", d["readgrid_points"], fmt="%.17g")
    (tmp_path / "grid.1.ele").write_text(writer_fixture("c4h4/read-grid/grid.1.ele"))
    monkeypatch.chdir(tmp_path)
    monkeypatch.setattr("sys.stdin", io.StringIO(READGRID_STREAM))
    print("case driver c4h4/read-grid from the keyword stream", flush=True)
    driver.main(["-"])
    out = (tmp_path / "jvec.vtu").read_text()
    ref = writer_fixture("c4h4/read-grid/reference/jvec.vtu")
    assert runtest_mismatches(out, ref, 1e-8) == []
