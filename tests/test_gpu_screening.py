"""GPU: screening block lists and spatial point order.

Rows are sorted along a Morton curve before evaluation and each contraction
row tile visits only the 32-function blocks that screening can leave
non-zero over it (basis.f90:118-136 zeroes a contraction beyond its
radius).  These tests check what that must not change: every point's
tensor whatever the caller's order (to summation-order rounding, 1e-12
relative), exact zeros where every block is screened, full lists when
screening is off, and the FP64 path against the oracle on a grid where the
lists leave most blocks out.
"""
import numpy as np
import pytest

from conftest import rel_err
from oracle import oracle, sph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cage():
    from gimic_amd import synth
    s = synth.cage60()
    ctx = s.context(screening=1e-8)
    yield s, ctx
    ctx.close()


def _grid(n, spacing, origin):
    from gimic_amd import synth
    return synth.cube_grid(n, spacing, 0.0) + np.asarray(origin, dtype=np.float64)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_result_does_not_depend_on_point_order(cage, precision):
    s, ctx = cage
    pts = _grid(24, 0.15, [3.0, -1.8, -1.8])  # across the cage wall
    perm = np.random.default_rng(3).permutation(len(pts))
    print("case cage60 point order", precision, flush=True)
    J = ctx.jtensor(pts, precision=precision)
    Jp = ctx.jtensor(pts[perm], precision=precision)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    e = rel_err(Jp[inv], J)
    assert e.max() <= 1e-12, e.max()
    # the same order twice gives the same bits
    assert np.array_equal(ctx.jtensor(pts, precision=precision), J)


def test_fp64_tiles_of_64_and_128_rows_agree(cage):
    """a chunk of up to 131,072 points takes 64-row FP64 tiles, a larger one
    128-row tiles: the same points evaluated alone and inside a larger batch
    give the same tensors to summation-order rounding"""
    s, ctx = cage
    pts = _grid(16, 0.2, [4.5, -1.5, -1.5])  # 4096 points next to the cage wall
    far = np.random.default_rng(8).uniform(-7.0, 7.0, (140000, 3))
    print("case cage60 fp64 tiles of 64 and 128 rows", flush=True)
    ctx.set_tile_stats(True)
    J64 = ctx.jtensor(pts, precision="fp64")
    assert ctx.tile_stats()["tile_rows"] == 64
    ctx.set_tile_stats(True)
    J128 = ctx.jtensor(np.concatenate([pts, far]), precision="fp64")[:len(pts)]
    assert ctx.tile_stats()["tile_rows"] == 128
    ctx.set_tile_stats(False)
    e = rel_err(J64, J128)
    assert e.max() <= 1e-12, e.max()


def test_chunks_change_launches_not_results(cage):
    """a chunk limit that splits the 24^3 points into four mixed launches
    (4096 rows of 1824 padded functions at 18 bytes) and six FP64 launches
    (2304 rows at 32 bytes) gives the tensors of one launch, to the rounding
    of each precision; 0 restores the default (8 GiB, one launch each), -1
    selects the large setting and smaller sizes are refused"""
    s, ctx = cage
    pts = _grid(24, 0.15, [3.0, -1.8, -1.8])
    assert ctx.chunk_bytes() == 8 << 30
    print("case cage60 chunk size", flush=True)
    radius = ctx.fp64_refined()[0]

    def launches(precision):
        # contraction launches of one call (the FP64 pass near nuclei off,
        # so only this precision's kernel is counted)
        ctx.set_fp64_radius(0.0)
        ctx.set_timing(True)
        try:
            J = ctx.jtensor(pts, precision=precision)
            n = ctx.get_timing()["contract"][1]
        finally:
            ctx.set_timing(False)
            ctx.set_fp64_radius(radius)
        return J, n

    ref = {}
    for p in ("fp64", "mixed"):
        ref[p], n = launches(p)
        assert n == 1, (p, n)
    small = 1824 * 18 * 4096
    try:
        ctx.set_chunk_bytes(small)
        assert ctx.chunk_bytes() == small
        j64, n64 = launches("fp64")
        jm, nm = launches("mixed")
    finally:
        ctx.set_chunk_bytes(0)
    assert (n64, nm) == (6, 4)
    assert ctx.chunk_bytes() == 8 << 30
    assert rel_err(j64, ref["fp64"]).max() <= 1e-12
    assert rel_err(jm, ref["mixed"]).max() <= 1e-5
    assert rel_err(jm, ref["fp64"]).max() <= 1e-5
    ctx.set_chunk_bytes(-1)
    assert ctx.chunk_bytes() >= 1 << 30
    ctx.set_chunk_bytes(0)
    with pytest.raises(RuntimeError):
        ctx.set_chunk_bytes(-2)


def test_lists_leave_out_blocks_and_fp64_matches_oracle(cage):
    s, ctx = cage
    pts = _grid(16, 0.2, [4.5, -1.5, -1.5])  # 4096 points near one side of the cage
    ctx.set_tile_stats(True)
    print("case cage60 block lists fp64", flush=True)
    J = ctx.jtensor(pts, precision="fp64")
    st = ctx.tile_stats()
    ctx.set_tile_stats(False)
    # a chunk of up to 131,072 points takes 64-row FP64 tiles
    assert st["tile_rows"] == 64 and st["tiles"] == len(pts) // 64
    frac = st["blocks"] / (st["tiles"] * st["nblk"])
    assert frac < 0.8, frac  # most of the far side's blocks are screened
    # the live groups are within the listed blocks: k <= 4 n on every tile,
    # so sum k^2 <= 16 sum n^2
    assert 0 < st["groups_sq"] <= 16 * st["blocks_sq"]
    sub = np.random.default_rng(5).choice(len(pts), 48, replace=False)
    dens = sph.to_cartesian(s.dens, s.shell_l)  # the oracle is cartesian: D_cart = T^T D T
    ref = oracle.jtensor(s.arrays(), dens, pts[sub], screening=1e-8)
    assert rel_err(J[sub], ref).max() <= 1e-10


def test_points_beyond_every_radius_give_zero(cage):
    s, ctx = cage
    pts = _grid(8, 0.5, [60.0, 60.0, 60.0])  # 512 points far outside every screening radius
    ctx.set_tile_stats(True)
    for precision in ("fp64", "mixed"):
        print("case far points", precision, flush=True)
        J = ctx.jtensor(pts, precision=precision)
        assert np.all(J == 0.0)
    st = ctx.tile_stats()
    ctx.set_tile_stats(False)
    assert st["tiles"] > 0 and st["blocks"] == 0 and st["groups_sq"] == 0


def test_unscreened_lists_are_full():
    from gimic_amd import synth
    s = synth.ring12()
    pts = _grid(8, 0.6, [-2.1, -2.1, -2.1])
    with s.context(screening=None) as ctx:
        ctx.set_tile_stats(True)
        J = ctx.jtensor(pts, precision="fp64")
        st = ctx.tile_stats()
        nbf = ctx.nbf
    assert st["blocks"] == st["tiles"] * st["nblk"]
    # every 8-function group holding a function is live on every tile
    assert st["groups_sq"] == st["tiles"] * ((nbf + 7) // 8) ** 2
    dens = sph.to_cartesian(s.dens, s.shell_l)
    ref = oracle.jtensor(s.arrays(), dens, pts[:40], screening=None)
    assert rel_err(J[:40], ref).max() <= 1e-10
