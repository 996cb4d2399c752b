"""The benchmark workloads at their stated shapes on the HIP path.

* Batched integration scan (sheet30, tetracene C18H12, nbf 312): 256 planes
  perpendicular to bond 1-2, offsets -6..+6 bohr, uneven order-7 Lobatto
  grids of 40..120 x 60..200 points, in one hipGraph launch.  FP64 tensors
  against the oracle on points drawn across the planes (1e-10 per point);
  one whole plane's FP64 flux against the oracle's quadrature (1e-10); every
  plane's mixed flux against FP64 (1e-6 of the plane's flux, floored at
  1e-3 of the largest); graph replays equal eager calls bit for bit.
* Open-shell spin currents (sheet200, C180H20, nbf 2620, alpha/beta nocc
  610/600, 8 matrices) on the 192 x 192 x 96 grid at 0.2 bohr: FP64 charge
  (total) and spin (spindens) tensors against the oracle on 24 points
  (1e-10 per point); mixed against FP64 over the whole grid (1e-5 per point,
  floor 1e-3 of the largest tensor), both cases from one pass.

* Precision (cage60, C60 truncated icosahedron, nbf 1800, nocc 180) on the
  128^3 grid at 0.12 bohr: mixed (with its FP64 pass near nuclei) against
  FP64 over all 2,097,152 points (1e-5 per point), and the flux through the
  bond plane of atom 1 and its nearest neighbour (140 x 140 order-7 Lobatto
  points, 14 x 14 bohr) within 1e-6 of the FP64 flux.

Per-point error: |dJ| / max(|J|, 1e-3 max |J|), as everywhere in the tests.
"""
import numpy as np
import pytest

from conftest import rel_err
from oracle import oracle, sph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scan():
    from gimic_amd import synth
    s = synth.sheet30()
    ctx = s.context(screening=1e-8)
    ctx.set_chunk_bytes(-1)
    planes = synth.scan_planes(s)
    yield s, ctx, planes
    ctx.close()


def _oracle_tensors(s, pts, spin="total"):
    return oracle.jtensor(s.arrays(), sph.to_cartesian(s.dens, s.shell_l), pts, screening=1e-8,
                          spin=spin)


def test_scan_shape(scan):
    s, ctx, planes = scan
    n = [g.size for g, _ in planes]
    assert s.natoms == 30 and s.nbf == 312 and len(planes) == 256
    assert all(g.npts[0] % 7 == 0 and g.npts[1] % 7 == 0 and g.npts[2] == 1 for g, _ in planes)
    assert 42 * 63 <= min(n) and max(n) <= 126 * 203
    assert 2_500_000 < sum(n) < 3_500_000


def test_scan_fp64_tensors_match_oracle(scan):
    """the scan's own tensors (host-array planes in graph mode, which return
    them) at 320 points spread over all 256 planes"""
    from gimic_amd import flux
    s, ctx, planes = scan
    print("case scan fp64 tensors vs oracle", flush=True)
    ctx.set_graph(True)
    try:
        _, J = ctx.flux_planes([flux.plane_of(g, b) for g, b in planes], precision="fp64",
                               return_jt=True)
    finally:
        ctx.set_graph(False)
    offs = np.cumsum([0] + [g.size for g, _ in planes])
    rng = np.random.default_rng(5)
    pick = np.concatenate([offs[k] + rng.choice(offs[k + 1] - offs[k], 1 + (k % 2), replace=False)
                           for k in range(256)])[:320]
    pts = np.concatenate([g.points() for g, _ in planes])[pick]
    Jo = _oracle_tensors(s, pts)
    e = rel_err(J[pick], Jo)
    assert e.max() <= 1e-10, e.max()


def test_scan_plane_flux_matches_oracle_quadrature(scan):
    """the smallest plane of the scan: the FP64 flux of the batched graph
    call against the oracle's tensors summed with the plane's weights"""
    from test_cpu_grid_flux import flux_sums
    s, ctx, planes = scan
    k = int(np.argmin([g.size for g, _ in planes]))
    g, B = planes[k]
    print(f"case scan plane {k} flux vs oracle quadrature ({g.size} points)", flush=True)
    ctx.set_graph(True)
    try:
        f = ctx.flux_plane_grids([p for p, _ in planes], [b for _, b in planes], precision="fp64")
    finally:
        ctx.set_graph(False)
    (net, pos, neg), _ = flux_sums(_oracle_tensors(s, g.points()), g, B)
    scale = abs(pos) + abs(neg)
    assert abs(f[k, 0] - net) <= 1e-10 * scale
    assert abs(f[k, 1] - pos) <= 1e-10 * scale and abs(f[k, 2] - neg) <= 1e-10 * scale


def test_scan_mixed_flux_and_graph_replays(scan):
    s, ctx, planes = scan
    grids, fields = [g for g, _ in planes], [b for _, b in planes]
    res = {}
    for prec in ("fp64", "mixed"):
        print(f"case scan {prec} eager and graph", flush=True)
        ctx.set_graph(False)
        eager = ctx.flux_plane_grids(grids, fields, precision=prec, modulus=True)
        ctx.set_graph(True)
        try:
            builds0 = ctx.graph_stats()[0]
            for rep in range(3):
                graph = ctx.flux_plane_grids(grids, fields, precision=prec, modulus=True)
                for a, b in zip(graph, eager):
                    assert np.array_equal(a, b), (prec, rep)
                # eager work between the replays
                ctx.jtensor(grids[0].points()[:100], precision=prec)
            assert ctx.graph_stats()[0] == builds0 + 1
        finally:
            ctx.set_graph(False)
        res[prec] = eager
    f64, fmx = res["fp64"][0][:, 0], res["mixed"][0][:, 0]
    floor = 1e-3 * np.abs(f64).max()
    rel = np.abs(fmx - f64) / np.maximum(np.abs(f64), floor)
    print("scan mixed flux error: max %.2e median %.2e" % (rel.max(), np.median(rel)))
    assert rel.max() <= 1e-6
    m64, mmx = res["fp64"][1][:, 0], res["mixed"][1][:, 0]
    relm = np.abs(mmx - m64) / np.maximum(np.abs(m64), 1e-3 * np.abs(m64).max())
    assert relm.max() <= 1e-6


@pytest.fixture(scope="module")
def sheet200():
    from gimic_amd import synth
    s = synth.sheet200()
    ctx = s.context(screening=1e-8)
    ctx.set_chunk_bytes(-1)
    yield s, ctx
    ctx.close()


def test_openshell_shape(sheet200):
    s, ctx = sheet200
    assert s.natoms == 200 and s.nbf == 2620 and s.dens.shape == (2, 4, 2620, 2620)
    assert ctx.info()["nset"] == 2


@pytest.mark.parametrize("spin", ["total", "spindens"])
def test_openshell_fp64_matches_oracle(sheet200, spin):
    s, ctx = sheet200
    rng = np.random.default_rng(9)
    # 16 points over the grid's box, 8 next to atoms (where the tight
    # functions and the GIAO terms are largest)
    pts = np.concatenate([rng.uniform([-19.1, -19.1, -9.5], [19.1, 19.1, 9.5], (16, 3)),
                          np.asarray(s.coords)[rng.choice(200, 8, replace=False)]
                          + rng.normal(0.0, 0.3, (8, 3))])
    print(f"case sheet200 fp64 {spin} vs oracle", flush=True)
    J = ctx.jtensor(pts, spin=spin, precision="fp64")
    Jo = _oracle_tensors(s, pts, spin=spin)
    e = rel_err(J, Jo)
    assert e.max() <= 1e-10, e.max()


def test_openshell_mixed_full_grid(sheet200):
    import torch
    s, ctx = sheet200
    ax = 0.2 * np.arange(192) - 19.1
    az = 0.2 * np.arange(96) - 9.5
    z, y, x = np.meshgrid(az, ax, ax, indexing="ij")
    pts = torch.tensor(np.stack([x.ravel(), y.ravel(), z.ravel()], 1), dtype=torch.float64,
                       device="cuda")
    assert pts.shape[0] == 192 * 192 * 96
    spins = ["total", "spindens"]
    print("case sheet200 192x192x96 fp64", flush=True)
    J64 = ctx.jtensor_multi_device(pts, spins, precision="fp64")
    print("case sheet200 192x192x96 mixed", flush=True)
    Jmx = ctx.jtensor_multi_device(pts, spins, precision="mixed")
    torch.cuda.synchronize()
    for k, name in enumerate(("charge", "spin")):
        nb = torch.linalg.norm(J64[k], dim=1)
        d = torch.linalg.norm(Jmx[k] - J64[k], dim=1)
        e = (d / torch.clamp(nb, min=1e-3 * float(nb.max()))).cpu().numpy()
        print("sheet200 %s per-point error max %.2e p99 %.2e" % (name, e.max(),
                                                                np.quantile(e, 0.99)))
        assert e.max() <= 1e-5, (name, e.max())
    # the two cases from one pass equal separate single-case evaluations
    Jt = ctx.jtensor_device(pts[:50000], spin="spindens", precision="fp64")
    assert rel_err(Jt.cpu().numpy(), J64[1, :50000].cpu().numpy()).max() <= 1e-12


def test_cage_mixed_full_grid_and_bond_plane_flux():
    import torch
    from gimic_amd import flux, synth
    from gimic_amd import grid as gridmod
    s = synth.cage60()
    ctx = s.context(screening=1e-8)
    try:
        ctx.set_chunk_bytes(-1)
        pts = torch.tensor(synth.cube_grid(128, 0.12, -7.62), dtype=torch.float64, device="cuda")
        assert pts.shape[0] == 128 ** 3
        print("case cage60 128^3 fp64", flush=True)
        J64 = ctx.jtensor_device(pts, precision="fp64")
        print("case cage60 128^3 mixed", flush=True)
        Jmx = ctx.jtensor_device(pts, precision="mixed")
        torch.cuda.synchronize()
        assert ctx.fp64_refined()[1] > 0  # the FP64 pass near nuclei ran
        nb = torch.linalg.norm(J64, dim=1)
        d = torch.linalg.norm(Jmx - J64, dim=1)
        e = (d / torch.clamp(nb, min=1e-3 * float(nb.max()))).cpu().numpy()
        print("cage60 per-point error max %.2e p99 %.2e" % (e.max(), np.quantile(e, 0.99)))
        assert e.max() <= 1e-5, e.max()
        del J64, Jmx
        c = np.asarray(s.coords)
        dist = np.linalg.norm(c - c[0], axis=1)
        dist[0] = np.inf
        j = int(np.argmin(dist))
        kw = {"_arg": "bond", "type": "lobatto", "gauss_order": 7, "bond": [1, j + 1],
              "fixcoord": [0.0, 0.0, 0.0], "distance": float(dist[j] / 2),
              "height": [-7.0, 7.0], "width": [-7.0, 7.0], "grid_points": [140, 140, 0]}
        g = gridmod.new_grid(kw, c)
        B = gridmod.get_magnet(g, "X")
        print("case cage60 bond plane flux", flush=True)
        f64 = np.array(flux.integrate_current(ctx, g, B, precision="fp64"))
        fmx = np.array(flux.integrate_current(ctx, g, B, precision="mixed"))
        rel = abs(fmx[0] - f64[0]) / abs(f64[0])
        print("cage60 bond plane flux %.6e, mixed rel error %.2e" % (f64[0], rel))
        assert abs(f64[0]) > 0.1 and rel <= 1e-6, (f64, fmx)
    finally:
        ctx.close()
