"""GPU parity: the HIP library (through its C ABI) against the oracle, the
flang-reference golden vectors and the reference's own fixtures.

Tolerances: FP64 path <= 1e-10 relative per point (error floor 1e-3 of the
largest tensor in the set); the mixed path (FP32 operands split into three BF16 parts) <= 1e-5 per point
with the same floor.
"""

import numpy as np
import pytest

from conftest import golden, rel_err
from oracle import molio, oracle

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
MIXED_TOL = 1e-5


@pytest.fixture(scope="module")
def ring12():
    from gimic_amd import synth
    s = synth.ring12()
    ctx = s.context(screening=1e-8)
    yield s, ctx
    ctx.close()


@pytest.fixture(scope="module")
def c4h4(c4h4_files):
    from gimic_amd.lib import Context
    mol, xd = c4h4_files
    ctx = Context.from_files(mol, xd, uhf=False, screening=1e-8)
    b = molio.read_mol(mol)
    dens = molio.read_xdens(xd, b)
    yield ctx, b, dens
    ctx.close()


@pytest.fixture(scope="module")
def openshell(openshell_files):
    from gimic_amd.lib import Context
    mol, xd = openshell_files
    ctx = Context.from_files(mol, xd, uhf=True, screening=1e-8)
    b = molio.read_mol(mol)
    dens = molio.read_xdens(xd, b, uhf=True)
    yield ctx, b, dens
    ctx.close()


def test_ring12_fp64_matches_flang_reference_all_components(ring12):
    """spherical basis on the GPU vs the reference fed D_cart = T^T D T"""
    s, ctx = ring12
    g = golden("ring12_ref.npz")
    print("case ring12 fp64 golden", flush=True)
    for flags in [(1, 1, 1), (0, 1, 1), (1, 0, 1), (1, 1, 0)]:
        ctx.set_flags(*flags)
        J = ctx.jtensor(g["points"], precision="fp64")
        e = rel_err(J, g["jt_%d%d%d" % flags])
        assert e.max() <= FP64_TOL, (flags, e.max())
    ctx.set_flags(True, True, True)


def test_ring12_unscreened_matches_reference():
    from gimic_amd import synth
    s = synth.ring12()
    g = golden("ring12_ref.npz")
    with s.context(screening=None) as ctx:
        J = ctx.jtensor(g["points"])
    assert rel_err(J, g["jt_noscreen"]).max() <= FP64_TOL


def test_c4h4_turbomole_fp64_matches_oracle(c4h4):
    ctx, b, dens = c4h4
    pts = np.random.default_rng(1).uniform(-7, 7, (300, 3))
    print("case c4h4 fp64 oracle", flush=True)
    J = ctx.jtensor(pts)
    Jo = oracle.jtensor(b.arrays(), dens, pts)
    assert rel_err(J, Jo).max() <= FP64_TOL


def test_c4h4_readgrid_fixture(c4h4):
    """test/c4h4/read-grid reference jvec.vtu (B = (0,0,-1), printed 3e20.10)"""
    ctx, _, _ = c4h4
    d = golden("c4h4.npz")
    print("case c4h4 read-grid fixture", flush=True)
    J = ctx.jtensor(d["readgrid_points"])
    jv = -J[:, 6:9]
    ref = d["readgrid_jvec"]
    assert np.all(np.abs(jv - ref) <= 5e-9 * np.abs(ref) + 1e-10 * np.abs(ref).max())


@pytest.mark.parametrize("spin", ["total", "alpha", "beta", "spindens"])
def test_openshell_spin_cases_match_oracle(openshell, spin):
    ctx, b, dens = openshell
    pts = np.random.default_rng(4).uniform(-6, 6, (200, 3))
    print(f"case open-shell {spin}", flush=True)
    J = ctx.jtensor(pts, spin=spin)
    Jo = oracle.jtensor(b.arrays(), dens, pts, spin=spin)
    assert rel_err(J, Jo).max() <= FP64_TOL


@pytest.mark.parametrize("npts", [0, 1, 127, 129, 1000])
def test_ragged_batch_sizes(ring12, npts):
    s, ctx = ring12
    g = golden("ring12_ref.npz")
    pts = np.resize(g["points"], (npts, 3)) if npts else np.zeros((0, 3))
    J = ctx.jtensor(pts)
    assert J.shape == (npts, 9)
    if npts:
        ref = np.resize(g["jt_111"], (npts, 9))
        assert rel_err(J, ref).max() <= FP64_TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_host_batch_in_sub_batches_matches_device_path(ring12, precision):
    """a host-array batch of more than 1,376,256 points runs in sub-batches
    (here 2 x 619,233, 262,144 and 65,536 points) whose copies overlap the
    evaluation of their neighbours: every point equals the one-launch device
    evaluation (FP64 to summation-order rounding, mixed to its tolerance)"""
    import torch
    s, ctx = ring12
    pts = np.random.default_rng(11).uniform(-6.0, 6.0, (1_566_146, 3))
    pts[::97] *= 0.05  # points next to the nuclei too (the FP64 pass)
    print("case ring12 host batch 1,566,146 points", precision, flush=True)
    J = ctx.jtensor(pts, precision=precision)
    d = torch.tensor(pts, dtype=torch.float64, device="cuda")
    Jd = ctx.jtensor_device(d, precision=precision).cpu().numpy()
    e = rel_err(J, Jd)
    assert e.max() <= (1e-12 if precision == "fp64" else 1e-5), e.max()
    if precision == "mixed":
        # the FP64 pass near nuclei counts the points of every sub-batch
        n_dev = ctx.fp64_refined()[1]
        ctx.jtensor(pts, precision=precision)
        assert ctx.fp64_refined()[1] == n_dev > 0
    # and repeats bit for bit
    assert np.array_equal(ctx.jtensor(pts, precision=precision), J)


def test_mixed_precision_within_stated_tolerance(ring12, c4h4):
    s, ctx = ring12
    g = golden("ring12_ref.npz")
    print("case mixed ring12", flush=True)
    Jm = ctx.jtensor(g["points"], precision="mixed")
    e = rel_err(Jm, g["jt_111"])
    print("ring12 mixed per-point error max %.3e p99 %.3e" % (e.max(), np.quantile(e, 0.99)))
    assert e.max() <= MIXED_TOL
    ctx4, b, dens = c4h4
    pts = np.random.default_rng(9).uniform(-7, 7, (500, 3))
    Jm = ctx4.jtensor(pts, precision="mixed")
    J = ctx4.jtensor(pts, precision="fp64")
    e = rel_err(Jm, J)
    print("c4h4 mixed per-point error max %.3e p99 %.3e" % (e.max(), np.quantile(e, 0.99)))
    assert e.max() <= MIXED_TOL


@pytest.mark.parametrize("variant,split", [(3, 0), (3, 1), (3, 2), (3, 3),
                                           (4, 0), (4, 1), (4, 3), (5, 0), (5, 2),
                                           (8, 0), (8, 1), (8, 2), (9, 0), (9, 1), (9, 3)])
def test_each_mixed_kernel_within_tolerance(c4h4, variant, split):
    """every mixed contraction kernel with several N-split counts (1000
    points = 8 row tiles of 128, 1300 = 11, so the XCD groups of row tiles
    are full and padded) on ragged points: per-point error vs FP64 within
    the stated bound"""
    ctx, b, dens = c4h4
    pts = np.random.default_rng(11).uniform(-7, 7, (1000 + 300 * (split % 2), 3))
    J = ctx.jtensor(pts, precision="fp64")
    print("case mixed kernel", variant, split, flush=True)
    ctx.set_mixed_kernel(variant, split)
    try:
        Jm = ctx.jtensor(pts, precision="mixed")
        assert ctx.tuning()[0] == variant
    finally:
        ctx.set_mixed_kernel(0)
    e = rel_err(Jm, J)
    print("kernel %d/%d per-point error max %.3e p99 %.3e" % (variant, split, e.max(),
                                                             np.quantile(e, 0.99)))
    assert e.max() <= MIXED_TOL


def test_mixed_n_tiles_split_any_way_agree_to_fp64_rounding():
    """the 256-row kernels (8: two LDS buffers, 9: three) with 1, 2, 4 and 8
    N-ranges per row tile on the 60-atom cage: an N-tile is always the same
    four live groups of its row tile, so its FP32 sums are the same bits in
    every configuration and only the FP64 order of adding N-tiles and ranges
    differs (a tile assigned twice, dropped or built from the wrong groups
    would differ by far more)"""
    from gimic_amd import synth
    s = synth.cage60()
    ctx = s.context(screening=1e-8)
    try:
        ax = np.linspace(-6.0, 6.0, 20)
        pts = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
        ctx.set_fp64_radius(0.0)
        out = {}
        for variant, split in [(9, 4), (9, 1), (9, 2), (9, 8), (8, 4), (8, 1)]:
            print("case n-tile split", variant, split, flush=True)
            ctx.set_mixed_kernel(variant, split)
            out[(variant, split)] = ctx.jtensor(pts, precision="mixed")
        ref = out[(9, 4)]
        scale = np.abs(ref).max()
        for k, J in out.items():
            d = np.abs(J - ref).max() / scale
            print(k, "max deviation %.2e" % d)
            assert d <= 1e-13, (k, d)
    finally:
        ctx.close()


def test_mixed_open_shell_two_sets(openshell):
    """alpha and beta sets in one mixed launch (pre-split kernel) against
    separate FP64 evaluations"""
    import torch
    ctx, b, dens = openshell
    pts = torch.tensor(np.random.default_rng(12).uniform(-8, 8, (777, 3)), dtype=torch.float64,
                       device="cuda")
    print("case mixed open shell", flush=True)
    ctx.set_mixed_kernel(3, 2)
    try:
        Jm = ctx.jtensor_multi_device(pts, ["alpha", "beta"], precision="mixed").cpu().numpy()
    finally:
        ctx.set_mixed_kernel(0)
    for k, spin in enumerate(["alpha", "beta"]):
        J = ctx.jtensor(pts.cpu().numpy(), spin=spin, precision="fp64")
        assert rel_err(Jm[k], J).max() <= MIXED_TOL


def test_removed_mixed_kernels_are_refused(ring12):
    from gimic_amd.lib import GimicError
    s, ctx = ring12
    for v in (1, 2, 6, 7, 10):
        with pytest.raises(GimicError):
            ctx.set_mixed_kernel(v, 0)


def test_beta_on_closed_shell_is_an_error(ring12):
    from gimic_amd.lib import GimicError
    s, ctx = ring12
    with pytest.raises(GimicError, match="not open-shell"):
        ctx.jtensor(np.zeros((4, 3)), spin="beta")


def test_error_inside_a_host_batch_in_sub_batches(ring12):
    """an error raised by the first sub-batch's evaluation (beta on a closed
    shell, found on the host) ends the copy thread and reaches the caller;
    the context then evaluates large host batches as before"""
    from gimic_amd.lib import GimicError
    s, ctx = ring12
    pts = np.random.default_rng(5).uniform(-6.0, 6.0, (1_400_000, 3))
    print("case ring12 host batch 1,400,000 points, beta on a closed shell", flush=True)
    with pytest.raises(GimicError, match="not open-shell"):
        ctx.jtensor(pts, spin="beta")
    J = ctx.jtensor(pts)
    g = golden("ring12_ref.npz")
    ref = ctx.jtensor(g["points"])
    assert rel_err(ref, g["jt_111"]).max() <= FP64_TOL
    assert np.isfinite(J).all() and np.abs(J).max() > 0


def test_drop_in_c_abi(c4h4_files):
    """gimic_init / set_magnet / set_spin / calc_jtensor / calc_jvector as the
    reference's C interface (gimic_interface.f90:26-151)"""
    from gimic_amd import lib
    L = lib.load()
    mol, xd = c4h4_files
    L.gimic_init(mol.encode(), xd.encode())
    b = molio.read_mol(mol)
    dens = molio.read_xdens(xd, b)
    pts = np.random.default_rng(5).uniform(-5, 5, (8, 3))
    Jo = oracle.jtensor(b.arrays(), dens, pts, screening=1e-6)  # C-ABI threshold
    B = np.array([0.3, -0.5, 0.8])
    L.gimic_set_magnet(lib.ptr(B))
    L.gimic_set_spin(b"total")
    for k in range(len(pts)):
        r = np.ascontiguousarray(pts[k])
        jt = np.zeros(9)
        jv = np.zeros(3)
        L.gimic_calc_jtensor(lib.ptr(r), lib.ptr(jt))
        L.gimic_calc_jvector(lib.ptr(r), lib.ptr(jv))
        assert rel_err(jt[None], Jo[k:k + 1], floor_frac=1.0).max() <= FP64_TOL
        assert np.allclose(jv, Jo[k].reshape(3, 3, order="F") @ B, rtol=1e-10, atol=1e-16)
    L.gimic_finalize()
