"""GPU plane flux: the driver end to end on the reference's integration
fixtures, batched planes against single-plane calls, and the mixed path's
flux error against FP64."""
import io

import numpy as np
import pytest

from conftest import golden, write_fixture
from test_cpu_grid_flux import flux_sums, stdout_blocks

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,uhf", [("c4h4.npz", False), ("openshell.npz", True)])
def test_driver_integral_matches_reference_fixture(tmp_path, name, uhf):
    from gimic_amd import driver, inp
    from oracle import molio, oracle
    d = golden(name)
    mol, xd = write_fixture(str(tmp_path), name)
    kw = inp.parse(str(d["integration_inp"]))
    kw["basis"], kw["xdens"] = mol, xd
    print(f"case driver {name}", flush=True)
    kw, ctx, g, B, uhf2 = driver.setup(kw)
    assert uhf2 == uhf
    buf = io.StringIO()
    res = driver.run_integral(kw, ctx, g, B, uhf, out=buf)
    text = str(d["integration_stdout"])
    cur = stdout_blocks(text, "Induced current (au)")
    mod = stdout_blocks(text, "Induced mod current (au)")
    spins = driver.spin_cases(uhf)
    for k, spin in enumerate(spins):
        assert np.allclose(res[spin]["current"], cur[k], atol=6e-7), spin
        if mod:
            assert np.allclose(res[spin]["modulus"], mod[k], atol=6e-7), spin
    # the printed blocks reproduce the reference's lines
    assert "%13.6f" % cur[0][0] in buf.getvalue()
    # and the FP64 GPU flux agrees with the oracle to 1e-10 relative
    b = molio.read_mol(mol)
    dens = molio.read_xdens(xd, b, uhf=uhf)
    J = oracle.jtensor(b.arrays(), dens, g.points(), spin="total", screening=1e-8)
    c, m = flux_sums(J, g, B)
    assert abs(res["total"]["current"][0] - c[0]) <= 1e-10 * abs(c[0])
    assert abs(res["total"]["modulus"][0] - m[0]) <= 1e-10 * abs(m[0])
    ctx.close()


def test_batched_planes_equal_single_plane_calls():
    from gimic_amd import flux, grid as gridmod, synth
    s = synth.ring12()
    ctx = s.context()
    coords = np.asarray(s.coords)
    grids, fields = [], []
    for k, dist in enumerate(np.linspace(-1.0, 2.5, 5)):
        kw = {"_arg": "bond", "type": "lobatto", "gauss_order": 7, "bond": [1, 2],
              "fixcoord": [0.0, 0.0, 0.0], "distance": float(dist), "height": [-4.0, 4.0],
              "width": [-2.2, 9.8], "grid_points": [21 + 7 * k, 28 + 7 * (k % 2), 0]}
        g = gridmod.new_grid(kw, coords)
        grids.append(g)
        fields.append(gridmod.get_magnet(g, "X"))
    print("case batched planes", flush=True)
    many = flux.integrate_planes(ctx, grids, fields)
    for k in range(len(grids)):
        one = flux.integrate_current(ctx, grids[k], fields[k])
        assert np.allclose(many[k], one, rtol=1e-12, atol=1e-15)
    ctx.close()


def test_mixed_precision_flux_error_reported(tmp_path):
    """ring12 plane (order-7 Lobatto on 20 x 30 blocks): mixed vs FP64 flux"""
    from gimic_amd import flux, grid as gridmod, synth
    s = synth.ring12()
    ctx = s.context()
    kw = {"_arg": "bond", "type": "lobatto", "gauss_order": 7, "bond": [1, 2],
          "fixcoord": [0.0, 0.0, 0.0], "distance": 1.3133, "height": [-4.0, 4.0],
          "width": [-2.2, 9.8], "grid_points": [140, 210, 0]}
    g = gridmod.new_grid(kw, np.asarray(s.coords))
    B = gridmod.get_magnet(g, "X")
    f64 = np.array(flux.integrate_current(ctx, g, B, precision="fp64"))
    mix = np.array(flux.integrate_current(ctx, g, B, precision="mixed"))
    rel = np.abs(mix - f64) / np.abs(f64)
    print("ring12 plane flux fp64", f64, "mixed", mix, "rel err", rel)
    assert rel[0] <= 1e-6
    ctx.close()


def test_acid_integral_matches_deviatoric_norm():
    """ACID integrand dT^2 (acid.f90:11-40) equals |dev(sym J)|_F^2 up to the
    reference's 0.3333333 constant; checked on the GPU tensors of a plane"""
    from gimic_amd import flux, grid as gridmod, synth
    s = synth.ring12()
    ctx = s.context()
    kw = {"_arg": "bond", "type": "gauss", "gauss_order": 9, "bond": [1, 2],
          "fixcoord": [0.0, 0.0, 0.0], "distance": 1.3133, "height": [-4.0, 4.0],
          "width": [-2.2, 9.8], "grid_points": [36, 54, 0], "radius": 5.0}
    g = gridmod.new_grid(kw, np.asarray(s.coords))
    B = gridmod.get_magnet(g, "X")
    plane = flux.plane_of(g, B)
    print("case acid", flush=True)
    cur, acid, jt = ctx.flux_planes([plane], acid=True, return_jt=True)
    J = jt.reshape(-1, 3, 3).transpose(0, 2, 1)       # J[p, m, b]
    S = 0.5 * (J + J.transpose(0, 2, 1))
    dev = S - np.einsum("pii->p", S)[:, None, None] * np.eye(3) / 3.0
    dt2 = np.sum(dev * dev, axis=(1, 2))
    w = np.where(np.linalg.norm(g.points() - g.center(), axis=1) > 5.0, 0.0, g.weights())
    ref = np.sum(w * dt2)
    assert abs(acid[0] - ref) <= 2e-7 * ref
    ctx.close()


def _uneven_planes(coords, shift=0.0):
    from gimic_amd import flux, grid as gridmod
    planes = []
    for k, dist in enumerate(np.linspace(-2.0, 3.0, 7)):
        kw = {"_arg": "bond", "type": "lobatto", "gauss_order": 7, "bond": [1, 2],
              "fixcoord": [0.0, 0.0, 0.0], "distance": float(dist + shift),
              "height": [-4.0, 4.0], "width": [-2.2, 9.8],
              "grid_points": [14 + 7 * k, 21 + 7 * ((3 * k) % 5), 0]}
        g = gridmod.new_grid(kw, coords)
        planes.append(flux.plane_of(g, gridmod.get_magnet(g, "X")))
    return planes


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_launch_equals_eager(precision):
    """7 planes of uneven size in one captured hipGraph: every replay (with
    eager work and synchronisation between replays) equals the eager batch
    bit for bit; new geometry of the same sizes reuses the graph"""
    from gimic_amd import synth
    s = synth.ring12()
    ctx = s.context()
    coords = np.asarray(s.coords)
    planes = _uneven_planes(coords)
    print("case graph", precision, flush=True)
    eager = ctx.flux_planes(planes, precision=precision, modulus=True, acid=True)
    ctx.set_graph(True)
    rng = np.random.default_rng(5)
    for rep in range(3):
        got = ctx.flux_planes(planes, precision=precision, modulus=True, acid=True)
        for g, e in zip(got, eager):
            assert np.array_equal(g, e), (rep, g, e)
        ctx.jtensor(rng.normal(0, 2, (300, 3)), precision=precision)
    assert ctx.graph_stats() == (1, 3)
    planes2 = _uneven_planes(coords, shift=0.37)
    got2 = ctx.flux_planes(planes2, precision=precision, modulus=True, acid=True)
    assert ctx.graph_stats() == (1, 4)
    ctx.set_graph(False)
    eager2 = ctx.flux_planes(planes2, precision=precision, modulus=True, acid=True)
    for g, e in zip(got2, eager2):
        assert np.array_equal(g, e)
    assert not np.array_equal(got2[0], eager[0])
    ctx.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_plane_grids_equal_host_arrays(precision):
    """points and weights generated on the device from the 1-D rules give the
    same sums, bit for bit, as the host-built arrays, eager and in a graph
    (replays, then new geometry of the same sizes); a grid with three layers
    in p3 counts as three planes whose sums add up to the layered grid's"""
    from gimic_amd import flux, grid as gridmod, synth
    s = synth.ring12()
    ctx = s.context()
    coords = np.asarray(s.coords)

    def grids_at(shift):
        gs, bs = [], []
        for k, dist in enumerate(np.linspace(-2.0, 3.0, 7)):
            kw = {"_arg": "bond", "type": "lobatto", "gauss_order": 7, "bond": [1, 2],
                  "fixcoord": [0.0, 0.0, 0.0], "distance": float(dist + shift),
                  "height": [-4.0, 4.0], "width": [-2.2, 9.8],
                  "grid_points": [14 + 7 * k, 21 + 7 * ((3 * k) % 5), 0]}
            g = gridmod.new_grid(kw, coords)
            gs.append(g)
            bs.append(gridmod.get_magnet(g, "X"))
        return gs, bs

    grids, fields = grids_at(0.0)
    host = [flux.plane_of(g, b) for g, b in zip(grids, fields)]
    print("case device plane grids eager", precision, flush=True)
    ref = ctx.flux_planes(host, precision=precision, modulus=True, acid=True)
    got = ctx.flux_plane_grids(grids, fields, precision=precision, modulus=True, acid=True)
    for g, e in zip(got, ref):
        assert np.array_equal(g, e), (g, e)
    print("case device plane grids graph", precision, flush=True)
    ctx.set_graph(True)
    rng = np.random.default_rng(11)
    b0, _ = ctx.graph_stats()
    for rep in range(3):
        got = ctx.flux_plane_grids(grids, fields, precision=precision, modulus=True, acid=True)
        for g, e in zip(got, ref):
            assert np.array_equal(g, e), (rep, g, e)
        ctx.jtensor(rng.normal(0, 2, (300, 3)), precision=precision)
    grids2, fields2 = grids_at(0.41)
    got2 = ctx.flux_plane_grids(grids2, fields2, precision=precision, modulus=True, acid=True)
    assert ctx.graph_stats()[0] == b0 + 1  # one capture for the three replays and the move
    ctx.set_graph(False)
    ref2 = ctx.flux_planes([flux.plane_of(g, b) for g, b in zip(grids2, fields2)],
                           precision=precision, modulus=True, acid=True)
    for g, e in zip(got2, ref2):
        assert np.array_equal(g, e)
    assert not np.array_equal(got2[0], ref[0])
    # three layers along v3
    g3 = grids[2]
    g3.pts[2] = np.array([-0.2, 0.0, 0.25])
    g3.wgt[2] = np.array([0.3, 0.5, 0.2])
    print("case device plane grids layered", precision, flush=True)
    lay = ctx.flux_plane_grids([g3], [fields[2]], precision=precision, modulus=True)
    one = ctx.flux_planes([flux.plane_of(g3, fields[2])], precision=precision, modulus=True)
    for g, e in zip(lay, one):
        assert np.allclose(g, e, rtol=1e-12, atol=1e-16), (g, e)
    ctx.close()


def test_plane_grid_descriptors_are_checked():
    """n1 or n2 below 1 is an argument error, reported and not launched; a
    one-point plane gives w n.(J B) at r = o + p1 v1 + p2 v2 + p3 v3"""
    from gimic_amd import synth
    from gimic_amd.lib import ptr
    s = synth.ring12()
    ctx = s.context()
    L = ctx._lib
    # every array handed to the library is held by a name for the call
    one = np.ones(1, dtype=np.int32)
    zero = np.zeros(1, dtype=np.int32)
    o = np.array([0.3, -0.2, 0.5])
    ax = np.eye(3).reshape(-1)
    p3w3 = np.array([0.0, 1.0])
    x1, w1 = np.array([0.25]), np.array([0.5])
    x2, w2 = np.array([-0.5]), np.array([2.0])
    nz = np.array([0.0, 0.0, 1.0])
    rad = np.array([1e10])
    out = np.zeros(3)

    def call(n1, n2):
        return L.gimic_flux_plane_grids(ctx._h, 1, ptr(n1), ptr(n2), ptr(o), ptr(ax), ptr(p3w3),
                                        ptr(x1), ptr(w1), ptr(x2), ptr(w2), ptr(nz), ptr(nz),
                                        ptr(o), ptr(rad), 0, 0, ptr(out), None, None)

    for n1, n2 in ((zero, one), (one, zero)):
        assert call(n1, n2) != 0
        assert b"n1, n2" in L.gimic_last_error()
    print("case single-point plane", flush=True)
    assert call(one, one) == 0
    r = o + np.array([0.25, -0.5, 0.0])
    J = ctx.jtensor(r[None], precision="fp64")[0].reshape(3, 3)  # [b, m]
    expect = (1.0 * 2.0) * 0.5 * (J[2] @ nz)  # w = (w3 w2) w1, n.(J B) with B = n = z
    assert np.isclose(out[0], expect, rtol=1e-13, atol=1e-300)
    assert out[1] + out[2] == out[0]
    ctx.close()
