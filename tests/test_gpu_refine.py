"""GPU: the FP64 pass of mixed evaluations over the points next to a nucleus.

Points farther than the radius from every nucleus keep the mixed result bit
for bit; the points within it carry FP64 results (equal to the FP64 path up
to summation order); the count matches a host-side selection; tensors of
several density sets, flux contributions, eager and graph launches all go
through the same pass."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _points_near_and_far(coords, rng, n_far=3000, n_near=400):
    far = rng.uniform(coords.min(0) - 3.0, coords.max(0) + 3.0, (n_far, 3))
    atoms = coords[rng.integers(0, len(coords), n_near)]
    near = atoms + rng.normal(0.0, 0.25, (n_near, 3))
    pts = np.concatenate([far, near])
    return pts[rng.permutation(len(pts))]


def _near_mask(pts, coords, radius):
    d2 = ((pts[:, None, :] - coords[None]) ** 2).sum(-1).min(1)
    return d2 <= radius * radius


def test_refined_points_are_fp64_and_the_rest_untouched():
    from gimic_amd import synth
    s = synth.ring12()
    ctx = s.context()
    coords = np.asarray(s.coords)
    pts = _points_near_and_far(coords, np.random.default_rng(21))
    print("case refine ring12 radius 0 / 0.4 / fp64", flush=True)
    ctx.set_fp64_radius(0.0)
    j0 = ctx.jtensor(pts, precision="mixed")
    assert ctx.fp64_refined() == (0.0, 0)
    ctx.set_fp64_radius(0.4)
    j1 = ctx.jtensor(pts, precision="mixed")
    j1b = ctx.jtensor(pts, precision="mixed")
    j64 = ctx.jtensor(pts, precision="fp64")
    near = _near_mask(pts, coords, 0.4)
    assert 50 < near.sum() < len(pts)
    assert ctx.fp64_refined() == (0.4, int(near.sum()))
    assert np.array_equal(j1, j1b)  # deterministic selection and order
    assert np.array_equal(j1[~near], j0[~near])
    scale = np.abs(j64[near]).max(axis=1, keepdims=True)
    assert np.all(np.abs(j1[near] - j64[near]) <= 1e-12 * scale)
    # the mixed results there were much further off
    assert np.abs(j0[near] - j64[near]).max() > 100 * np.abs(j1[near] - j64[near]).max()
    with pytest.raises(RuntimeError):
        ctx.set_fp64_radius(-1.0)
    ctx.close()


def test_refinement_covers_every_density_set(openshell_files):
    from gimic_amd.lib import Context
    import torch
    mol, xd = openshell_files
    ctx = Context.from_files(mol, xd, uhf=True)
    coords = ctx.atoms()[0]
    pts = _points_near_and_far(coords, np.random.default_rng(4), 2000, 300)
    near = _near_mask(pts, coords, 0.4)
    d = torch.tensor(pts, device="cuda:0")
    spins = ["alpha", "beta", "spindens"]
    print("case refine openshell 3 sets", flush=True)
    ref = ctx.jtensor_multi_device(d, spins, precision="fp64").cpu().numpy()
    ctx.set_fp64_radius(0.0)
    m0 = ctx.jtensor_multi_device(d, spins, precision="mixed").cpu().numpy()
    ctx.set_fp64_radius(0.4)
    m1 = ctx.jtensor_multi_device(d, spins, precision="mixed").cpu().numpy()
    torch.cuda.synchronize()
    assert ctx.fp64_refined()[1] == int(near.sum()) > 0
    for k in range(len(spins)):
        assert np.array_equal(m1[k][~near], m0[k][~near])
        scale = np.abs(ref[k][near]).max(axis=1, keepdims=True) + 1e-300
        assert np.all(np.abs(m1[k][near] - ref[k][near]) <= 1e-12 * scale), spins[k]
    ctx.close()


def _plane_through_atom(s):
    from gimic_amd import grid as gridmod
    c = np.asarray(s.coords)
    kw = {"_arg": "bond", "type": "lobatto", "gauss_order": 7, "bond": [1, 2],
          "fixcoord": [0.0, 0.0, 0.0], "distance": 0.02, "height": [-4.0, 4.0],
          "width": [-2.2, 9.8], "grid_points": [70, 84, 0]}
    g = gridmod.new_grid(kw, c)
    return g, gridmod.get_magnet(g, "X")


def test_flux_with_refinement_eager_equals_graph_and_matches_its_tensors():
    """a plane 0.02 bohr from atom 1: the refined flux is the quadrature of
    the refined tensors (the contributions were replaced with the tensors),
    those tensors are FP64 next to the nucleus, and the graph launch equals
    the eager call bit for bit"""
    from gimic_amd import flux, synth
    s = synth.ring12()
    ctx = s.context()
    coords = np.asarray(s.coords)
    g, B = _plane_through_atom(s)
    pl = flux.plane_of(g, B)
    near = _near_mask(pl["points"], coords, 0.4)
    assert near.sum() > 10
    print("case refine flux", flush=True)
    _, j64 = ctx.flux_planes([pl], precision="fp64", return_jt=True)
    f1, j1 = ctx.flux_planes([pl], precision="mixed", return_jt=True)
    assert ctx.fp64_refined()[1] == int(near.sum())
    scale = np.abs(j64[near]).max(axis=1, keepdims=True)
    assert np.all(np.abs(j1[near] - j64[near]) <= 1e-12 * scale)
    # host quadrature of the returned tensors (finalize's formula)
    J = j1.reshape(-1, 3, 3)  # [p, b, m]
    jv = np.einsum("pbm,b->pm", J, np.asarray(B))
    r = np.linalg.norm(pl["points"] - pl["center"], axis=1)
    c = np.where(r > pl["radius"], 0.0, pl["weights"] * (jv @ pl["normal"]))
    tot = np.abs(c).sum()
    assert abs(c.sum() - f1[0, 0]) <= 1e-13 * tot
    assert abs(c[c > 0].sum() - f1[0, 1]) <= 1e-13 * tot
    f1g = ctx.flux_plane_grids([g], [B], precision="mixed", modulus=True, acid=True)
    ctx.set_graph(True)
    for rep in range(2):
        fg = ctx.flux_plane_grids([g], [B], precision="mixed", modulus=True, acid=True)
        for a, b in zip(fg, f1g):
            assert np.array_equal(a, b), rep
        ctx.jtensor(np.zeros((5, 3)) + 1.0, precision="mixed")
    ctx.set_graph(False)
    assert np.array_equal(f1g[0], f1)
    ctx.close()


def test_radius_covering_everything_gives_fp64_and_zero_radius_nothing():
    from gimic_amd import synth
    s = synth.ring12()
    ctx = s.context()
    pts = _points_near_and_far(np.asarray(s.coords), np.random.default_rng(8), 700, 100)
    print("case refine radius 1e3", flush=True)
    j64 = ctx.jtensor(pts, precision="fp64")
    ctx.set_fp64_radius(1e3)
    jall = ctx.jtensor(pts, precision="mixed")
    assert ctx.fp64_refined() == (1e3, len(pts))
    scale = np.abs(j64).max(axis=1, keepdims=True) + 1e-300
    assert np.all(np.abs(jall - j64) <= 1e-12 * scale)
    # points far from every nucleus: nothing selected
    ctx.set_fp64_radius(0.4)
    far = np.asarray(s.coords).max(0) + 20.0 + np.random.default_rng(2).uniform(0, 1, (50, 3))
    ctx.jtensor(far, precision="mixed")
    assert ctx.fp64_refined() == (0.4, 0)
    ctx.close()
