"""Pins the surrogate's CPU oracle (oracle/mlp_ref.py) by mathematics:
central finite differences of its FP64 loss, the SiLU derivative, the loss
against a direct per-point sum, and bfloat16 rounding on known values."""
import numpy as np
import pytest

from oracle import mlp_ref as M


def _net(seed=0, sigma=1.5):
    rng = np.random.default_rng(seed)
    lay, n = M.layout()
    p = np.zeros(n)
    for w, b, kin, nout in lay:
        p[w:b] = rng.normal(0, np.sqrt(2.0 / kin), kin * nout)
        p[b:b + nout] = rng.normal(0, 0.1, nout)
    freqs = rng.normal(0, sigma, (M.F, 3))
    cs = np.array([0.3, -0.2, 0.1, 4.0])
    os_ = rng.uniform(0.5, 2.0, 9)
    return p, freqs, cs, os_, rng


def test_layout():
    lay, n = M.layout()
    assert n == 160 * 128 + 128 + 3 * (128 * 128 + 128) + 128 * 16 + 16 == 72208
    assert lay[0][:2] == (0, 160 * 128)
    assert lay[-1][3] == 16


def test_dsilu_closed_form():
    z = np.linspace(-8, 8, 101)
    h = 1e-6
    fd = (M.silu(z + h) - M.silu(z - h)) / (2 * h)
    assert np.abs(fd - M.dsilu(z)).max() < 1e-8


def test_bf16_rounding():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert M.bf16(1.0 + 2.0 ** -8) == 1.0
    assert M.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert M.bf16(-3.0) == -3.0
    x = np.random.default_rng(1).normal(size=1000)
    assert np.abs(M.bf16(x) - x).max() <= np.abs(x).max() * 2.0 ** -8


def test_loss_direct_sum():
    p, freqs, cs, os_, rng = _net()
    r = rng.normal(0, 3, (17, 3))
    t = rng.normal(0, 1, (17, 9))
    eps = 0.05
    loss, _ = M.loss_grad(p, freqs, cs, os_, r, t, eps)
    J = M.predict(p, freqs, cs, os_, r)
    y, _ = M.forward(p, freqs, cs, r)
    direct = 0.0
    for i in range(17):
        tt = t[i] / os_
        Ji = y[i, :9] * np.exp(y[i, 9]) * os_
        assert np.allclose(Ji, J[i], rtol=1e-14)
        n = np.sqrt((tt ** 2).sum() + eps * eps)
        direct += ((y[i, :9] - tt / n) ** 2).sum() + (y[i, 9] - np.log(n)) ** 2
    assert abs(loss - direct / 17) < 1e-12 * abs(loss)


def test_gradient_finite_differences():
    p, freqs, cs, os_, rng = _net(seed=3)
    r = rng.normal(0, 3, (9, 3))
    t = rng.normal(0, 1, (9, 9))
    eps = 0.1
    _, g = M.loss_grad(p, freqs, cs, os_, r, t, eps)
    lay, _ = M.layout()
    # a few entries of every weight matrix and bias vector
    idx = []
    for w, b, kin, nout in lay:
        idx += list(rng.integers(w, b, 6)) + list(rng.integers(b, b + min(nout, 10), 3))
    # the log-amplitude row of the output layer
    w4, b4, kin4, _ = lay[-1]
    idx += list(rng.integers(w4 + 9 * kin4, w4 + 10 * kin4, 4)) + [b4 + 9]
    h = 1e-5
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_grad(pp, freqs, cs, os_, r, t, eps)[0] -
              M.loss_grad(pm, freqs, cs, os_, r, t, eps)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_unused_outputs_have_zero_gradient():
    p, freqs, cs, os_, rng = _net(seed=5)
    _, g = M.loss_grad(p, freqs, cs, os_, rng.normal(0, 3, (5, 3)), rng.normal(0, 1, (5, 9)), 0.1)
    w, b, kin, nout = M.layout()[0][-1]
    # rows 0..8 tensor elements, row 9 the log amplitude, rows 10..15 unused
    assert np.all(g[w + 10 * kin:b] == 0) and np.all(g[b + 10:b + nout] == 0)
    assert np.any(g[w + 9 * kin:w + 10 * kin] != 0) and g[b + 9] != 0


def test_bf16_mode_close_to_fp64():
    p, freqs, cs, os_, rng = _net(seed=7)
    r = rng.normal(0, 3, (64, 3))
    t = rng.normal(0, 1, (64, 9))
    l0, g0 = M.loss_grad(p, freqs, cs, os_, r, t, 0.1)
    l1, g1 = M.loss_grad(p, freqs, cs, os_, r, t, 0.1, bf=True)
    assert abs(l1 - l0) < 3e-2 * l0
    assert np.linalg.norm(g1 - g0) < 3e-2 * np.linalg.norm(g0)
    assert np.linalg.norm(g1 - g0) > 0  # the rounding is really applied


@pytest.mark.parametrize("seed", [11])
def test_features_periodicity(seed):
    # sin/cos features: a shift of r by scale * v with B v integer leaves
    # them unchanged (checks the 2 pi and the scaling)
    p, freqs, cs, os_, rng = _net(seed)
    f = np.zeros((M.F, 3))
    f[:, 0] = rng.integers(-3, 4, M.F)
    r = rng.normal(0, 2, (4, 3))
    x0 = M.features(r, f, cs)
    x1 = M.features(r + np.array([cs[3], 0, 0]), f, cs)
    assert np.abs(x0[:, :2 * M.F] - x1[:, :2 * M.F]).max() < 1e-12


@pytest.mark.parametrize("kind,layer", [("no_dsilu", l) for l in range(1, M.D + 1)] +
                         [("no_bias", l) for l in range(M.D + 1)])
def test_gpu_gradient_bounds_catch_a_planted_mistake(kind, layer):
    """a missing SiLU' or bias gradient in any one layer moves some weight or
    bias block by far more than the per-block bound of the GPU gradient test
    (so that test would fail on such a kernel)"""
    p, freqs, cs, os_, rng = _net(seed=13, sigma=2.0)
    r = rng.uniform(-6, 6, (512, 3))
    t = rng.normal(0, 1, (512, 9))
    _, g = M.loss_grad(p, freqs, cs, os_, r, t, 0.01, bf=True)
    _, gp = M.loss_grad(p, freqs, cs, os_, r, t, 0.01, bf=True, plant=(kind, layer))
    dev = max(np.linalg.norm(gp[lo:hi] - g[lo:hi]) / np.linalg.norm(g[lo:hi])
              for _, (lo, hi) in M.blocks())
    assert dev > 50 * M.GPU_BLOCK_TOL
