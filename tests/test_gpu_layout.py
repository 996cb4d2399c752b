"""The MFMA lane layouts the contraction kernels are written against.

With A = lane+1 and B = 1 (then the reverse) every output slot (lane l,
element r) reports the sum of the A lanes (B lanes) that feed it.  For both
builtins the A lanes feeding row i are i, i+16, i+32, i+48 and the B lanes
feeding column j are j, j+16, j+32, j+48 (16x16x4 f64: one value per lane;
16x16x32 bf16: 8 values per lane), so the sums decode to row and column.
"""
import numpy as np
import pytest

from gimic_amd import lib

pytestmark = pytest.mark.gpu


def probe():
    f64 = np.zeros(512)
    bf = np.zeros(512, dtype=np.float32)
    lib.check(lib.load().gimic_selftest_mfma_layout(lib.ptr(f64), lib.ptr(bf)))
    return f64, bf.astype(np.float64)


def decode(vals, per_lane):
    # sum over the 4 lane groups of (i + 16 g + 1), times values per lane
    return (vals / per_lane - 100.0) / 4.0


def kernel_row(l, r, kind):
    """row of output slot (lane l, element r) as the kernels assume"""
    g = l // 16
    return g + 4 * r if kind == "f64" else 4 * g + r


def test_mfma_layouts_match_kernel_assumptions():
    f64, bf = probe()
    for kind, vals, per in (("f64", f64, 1.0), ("bf16", bf, 8.0)):
        rows = decode(vals[:256], per).reshape(64, 4)
        cols = decode(vals[256:], per).reshape(64, 4)
        print(kind, "rows of lanes 0,16,32,48:", rows[[0, 16, 32, 48]].tolist())
        print(kind, "cols of lanes 0..3:", cols[:4].tolist())
        for l in range(64):
            for r in range(4):
                assert cols[l, r] == l % 16, (kind, l, r, cols[l, r])
                assert rows[l, r] == kernel_row(l, r, kind), (kind, l, r, rows[l, r])


def row32(l, e):
    """row of output slot (lane l, element e) of 32x32x16 BF16 as the
    32x32-block kernel assumes: 8 (e // 4) + 4 (l // 32) + e % 4"""
    return 8 * (e // 4) + 4 * (l // 32) + e % 4


def test_mfma32_layout_matches_kernel_assumptions():
    """32x32x16 BF16: A lanes feeding row i are i and i+32 (8 k each), B
    lanes feeding column j are j and j+32"""
    out = np.zeros(2048, dtype=np.float32)
    lib.check(lib.load().gimic_selftest_mfma32_layout(lib.ptr(out)))
    out = out.astype(np.float64)
    rows = ((out[:1024] / 8.0 - 34.0) / 2.0).reshape(64, 16)
    cols = ((out[1024:] / 8.0 - 34.0) / 2.0).reshape(64, 16)
    print("32x32 rows of lanes 0, 32:", rows[[0, 32]].tolist())
    print("32x32 cols of lanes 0..2:", cols[:3].tolist())
    for l in range(64):
        for e in range(16):
            assert cols[l, e] == l % 32, (l, e, cols[l, e])
            assert rows[l, e] == row32(l, e), (l, e, rows[l, e])
