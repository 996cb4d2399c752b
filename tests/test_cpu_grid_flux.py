"""Grid construction and plane-flux arithmetic against the reference's
integration fixtures (test/c4h4/integration, test/open-shell/integration),
with the tensors from the CPU oracle: pins the bond-grid geometry, the
quadrature, the field choice and the +/- flux split on the CPU."""
import re

import numpy as np
import pytest

from conftest import golden, stdout_setup_lines, write_fixture
from gimic_amd import grid as gridmod, inp
from oracle import molio, oracle

NUM = r"[-+]?\d+\.\d+"


def stdout_geometry(text):
    geo = {}
    for key in ("center", "origin", "basv1", "basv2", "basv3", "lenghts", "magnet"):
        m = re.search(rf"^ *{key} +({NUM}) +({NUM}) +({NUM})", text, re.M)
        geo[key] = np.array([float(m.group(i)) for i in (1, 2, 3)])
    return geo


def stdout_blocks(text, label):
    """(net, pos, neg) of every block whose first line contains `label`"""
    out = []
    lines = text.splitlines()
    for i, l in enumerate(lines):
        if label in l:
            net = float(l.split(":")[1])
            pos = float(lines[i + 1].split(":")[1].split("(")[0])
            neg = float(lines[i + 2].split(":")[1].split("(")[0])
            out.append((net, pos, neg))
    return out


def fixture_setup(tmp_path, name, uhf):
    d = golden(name)
    mol, xd = write_fixture(str(tmp_path), name)
    b = molio.read_mol(mol)
    dens = molio.read_xdens(xd, b, uhf=uhf)
    kw = inp.with_defaults(inp.parse(str(d["integration_inp"])))
    g = gridmod.new_grid(dict(kw["Grid"]), np.asarray(b.coords))
    B = gridmod.get_magnet(g, kw.get("magnet_axis", ""), kw.get("magnet"))
    return d, b, dens, kw, g, B


def flux_sums(J, g, B):
    """integral.f90:119-143 and 258-283 on precomputed tensors"""
    jv = np.einsum("pbm,b->pm", J.reshape(-1, 3, 3), B)
    n = g.normal()
    dot = jv @ n
    w = g.weights()
    r = np.linalg.norm(g.points() - g.center(), axis=1)
    w = np.where(r > g.radius, 0.0, w)
    c = dot * w
    sgn = np.where(np.abs(dot) < 1e-12, 0.0, np.sign(dot))
    cm = sgn * np.linalg.norm(jv, axis=1) * w
    split = lambda v: (v.sum(), v[v > 0].sum(), v[v <= 0].sum())
    return split(c), split(cm)


@pytest.mark.parametrize("name,uhf", [("c4h4.npz", False), ("openshell.npz", True)])
def test_bond_grid_geometry_matches_reference_printout(tmp_path, name, uhf):
    d, b, dens, kw, g, B = fixture_setup(tmp_path, name, uhf)
    geo = stdout_geometry(str(d["integration_stdout"]))
    tol = 5.01e-7  # f12.6
    assert np.allclose(g.bond_center, geo["center"], atol=tol)
    assert np.allclose(g.l, geo["lenghts"], atol=tol)
    for k in range(3):
        assert np.allclose(g.basv[:, k], geo[f"basv{k + 1}"], atol=tol)
    assert np.allclose(B, geo["magnet"], atol=tol)
    # printed origin is before the (zero) rotation and the quadrature
    m = re.search(r"Adjusted number of grid points for quadrature: +(\d+) +(\d+) +(\d+)",
                  str(d["integration_stdout"]))
    assert [int(m.group(1)), int(m.group(2))] == list(g.npts[:2])


@pytest.mark.parametrize("name,uhf", [("c4h4.npz", False), ("openshell.npz", True)])
def test_setup_lines_match_reference_printout(tmp_path, name, uhf):
    """new_grid's and check_field's lines as the reference prints them
    (c4h4: a zero rotation; both: the 'X' axis, no field check)"""
    d, b, dens, kw, g, B = fixture_setup(tmp_path, name, uhf)
    mlog = []
    gridmod.get_magnet(g, kw.get("magnet_axis", ""), kw.get("magnet"), log=mlog)
    want = stdout_setup_lines(str(d["integration_stdout"]))
    assert g.log + [" *** Grid plot in grid.xyz"] + mlog == want


@pytest.mark.parametrize("name,uhf", [("c4h4.npz", False), ("openshell.npz", True)])
def test_integral_printout_layout(tmp_path, name, uhf):
    """driver.integral_report rebuilt from the printed values: the open-shell
    run's four spin cases with |J|, byte for byte around the numbers"""
    from conftest import printout_mismatches, stdout_integral_lines
    from gimic_amd import driver
    d, b, dens, kw, g, B = fixture_setup(tmp_path, name, uhf)
    text = str(d["integration_stdout"])
    cur = stdout_blocks(text, "Induced current (au)")
    mod = stdout_blocks(text, "Induced mod current (au)")
    res = {spin: dict(current=cur[k], modulus=mod[k], acid=None)
           for k, spin in enumerate(driver.spin_cases(uhf))}
    # c4h4's printout is from a version that integrated |J| unasked
    jmod = bool(kw["Essential"]["jmod"]) or name == "c4h4.npz"
    mlog = []
    gridmod.get_magnet(g, kw.get("magnet_axis", ""), kw.get("magnet"), log=mlog)
    ours = driver.integral_report(res, uhf, jmod, bool(kw["Essential"]["acid"]), g.radius, mlog)
    assert printout_mismatches(ours, stdout_integral_lines(text), 1.5e-5) == []


@pytest.mark.parametrize("name,uhf", [("c4h4.npz", False), ("openshell.npz", True)])
def test_oracle_flux_matches_reference_integration_fixture(tmp_path, name, uhf):
    d, b, dens, kw, g, B = fixture_setup(tmp_path, name, uhf)
    text = str(d["integration_stdout"])
    cur = stdout_blocks(text, "Induced current (au)")
    mod = stdout_blocks(text, "Induced mod current (au)")
    spins = ["total", "alpha", "beta", "spindens"] if uhf else ["total"]
    pts = g.points()
    thrs = kw["Advanced"]["screening_thrs"] if kw["Advanced"]["screening"] else None
    for k, spin in enumerate(spins):
        J = oracle.jtensor(b.arrays(), dens, pts, spin=spin, screening=thrs)
        c, m = flux_sums(J, g, B)
        assert np.allclose(c, cur[k], atol=6e-7), (spin, c, cur[k])  # f13.6
        if mod:
            assert np.allclose(m, mod[k], atol=6e-7), (spin, m, mod[k])


def test_au2si_conversion_factor():
    from gimic_amd import flux
    assert f"{flux.au2si(1.0):13.6f}".strip() == "28.179409"


def test_inp_parser_reads_fixture_keywords():
    d = golden("c4h4.npz")
    kw = inp.parse(str(d["integration_inp"]))
    assert kw["calc"] == "integral" and kw["magnet_axis"] == "z"
    assert kw["Grid"]["_arg"] == "bond" and kw["Grid"]["gauss_order"] == 9
    assert kw["Grid"]["bond"] == [2, 1] and kw["Grid"]["width"] == [-1.25614, 6.0]
    assert kw["Advanced"]["screening_thrs"] == 1e-8 and kw["Advanced"]["spherical"] is False


def test_plane_grid_generator_rounding_matches_grid_points():
    """the device generator's arithmetic, ((o + p1 v1) + p2 v2) + p3 v3 and
    (w3 w2) w1 each rounded once per operation, restated in numpy, equals
    Grid.points() and Grid.weights() bit for bit on an uneven bond plane"""
    from gimic_amd import grid as gridmod
    coords = np.array([[0.0, 0.0, 0.0], [1.39, 0.21, -0.05], [2.1, 1.4, 0.3]])
    kw = {"_arg": "bond", "type": "lobatto", "gauss_order": 7, "bond": [1, 2],
          "fixcoord": [0.3, -0.7, 2.0], "distance": 0.37, "height": [-4.0, 4.5],
          "width": [-2.2, 9.8], "grid_points": [33, 47, 0]}
    g = gridmod.new_grid(kw, coords)
    p1, p2, p3 = g.pts
    w1, w2, w3 = g.wgt
    n1, n2 = len(p1), len(p2)
    idx = np.arange(n1 * n2 * len(p3))
    i, j, k = idx % n1, (idx // n1) % n2, idx // (n1 * n2)
    v = g.basv
    r = np.empty((len(idx), 3))
    for d in range(3):
        x = g.origin[d] + p1[i] * v[d, 0]
        x = x + p2[j] * v[d, 1]
        r[:, d] = x + p3[k] * v[d, 2]
    w = (w3[k] * w2[j]) * w1[i]
    assert np.array_equal(r, g.points())
    assert np.array_equal(w, g.weights())


def test_plane_grid_descriptors_match_each_grid():
    """the vectorised descriptors equal, bit for bit, each grid's own
    origin, axes, normal, centre and 1-D rules, layered grids included"""
    from gimic_amd import grid as gridmod
    from gimic_amd.lib import plane_grid_descriptors
    coords = np.array([[0.0, 0.0, 0.0], [1.39, 0.21, -0.05], [2.1, 1.4, 0.3]])
    grids, fields = [], []
    for k, dist in enumerate(np.linspace(-1.5, 2.5, 6)):
        kw = {"_arg": "bond", "type": "lobatto", "gauss_order": 7, "bond": [1, 2],
              "fixcoord": [0.3, -0.7, 2.0], "distance": float(dist), "height": [-4.0, 4.5],
              "width": [-2.2, 9.8], "grid_points": [14 + 7 * k, 21 + 7 * (k % 3), 0],
              "radius": 3.0 + k}
        g = gridmod.new_grid(kw, coords)
        grids.append(g)
        fields.append(gridmod.get_magnet(g, "X"))
    # one layer each (the direct path), then a layered grid (planes per layer)
    d = plane_grid_descriptors(grids, fields)
    for gi, g in enumerate(grids):
        assert np.array_equal(d["centers"][gi], g.center())
        assert np.array_equal(d["normals"][gi], g.normal())
        assert np.array_equal(d["axes"][gi], g.basv.T.reshape(-1))
        assert tuple(d["p3w3"][gi]) == (g.pts[2][0], g.wgt[2][0])
    assert np.array_equal(d["nodes2"], np.concatenate([g.pts[1] for g in grids]))
    assert list(d["n1"]) == [len(g.pts[0]) for g in grids]
    grids[3].pts[2] = np.array([-0.2, 0.0, 0.25])
    grids[3].wgt[2] = np.array([0.3, 0.5, 0.2])
    d = plane_grid_descriptors(grids, fields)
    rows = [(gi, k) for gi, g in enumerate(grids) for k in range(len(g.pts[2]))]
    assert len(d["n1"]) == len(rows) == 8
    o1 = o2 = 0
    for r, (gi, k) in enumerate(rows):
        g = grids[gi]
        assert d["owner"][r] == gi
        assert d["n1"][r] == len(g.pts[0]) and d["n2"][r] == len(g.pts[1])
        assert np.array_equal(d["origins"][r], g.origin)
        assert np.array_equal(d["axes"][r], g.basv.T.reshape(-1))
        assert np.array_equal(d["normals"][r], g.normal())
        assert np.array_equal(d["centers"][r], g.center())
        assert np.array_equal(d["bfields"][r], np.asarray(fields[gi], dtype=np.float64))
        assert d["radii"][r] == g.radius
        assert tuple(d["p3w3"][r]) == (g.pts[2][k], g.wgt[2][k])
        assert np.array_equal(d["nodes1"][o1:o1 + d["n1"][r]], g.pts[0])
        assert np.array_equal(d["wts2"][o2:o2 + d["n2"][r]], g.wgt[1])
        o1 += d["n1"][r]
        o2 += d["n2"][r]
    assert o1 == len(d["nodes1"]) and o2 == len(d["wts2"])
