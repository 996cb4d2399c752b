"""CPU oracle of the learned J-field surrogate (test infrastructure only).

The surrogate has no counterpart in the reference: the network is defined by
the library's gimic_mlp_* entry points and this module restates it in numpy
FP64 so the GPU forward pass, loss and parameter gradients can be checked.

Network (parameter layout identical to the library's flat vector, per layer
l: W_l[n_out][k_in] row-major, then b_l[n_out]):
    u  = (r - centre) / scale
    x  = [sin(2 pi B u) (64), cos(2 pi B u) (64), u (3), 0 (29)]   k_in 160
    h_l = SiLU(W_l h_{l-1} + b_l), l = 0..3                          width 128
    y  = W_4 h_3 + b_4                     16 rows: 9 used + log amplitude y_9
    J  = y[:9] * exp(min(y_9, 60)) * out_scale
Loss, with t = J_target / out_scale, n = sqrt(|t|^2 + eps^2), u = t / n:
    mean_p [ sum_i (y_i - u_i)^2 + (y_9 - log n)^2 ]

`bf16=True` rounds every GEMM operand to bfloat16 (round to nearest even)
where the device kernels do: features, weights, hidden activations, dY and
dZ for the products; bias gradients use the unrounded dY / dZ.

Parity pinned by: finite differences of this FP64 oracle (tests), the SiLU
derivative against its closed form, and the loss against a direct sum.
"""
import numpy as np

F, K0, H, D, NO, NOUT = 64, 160, 128, 4, 16, 9

# Bounds the GPU gradient test applies against the BF16-emulating mode of
# this oracle: whole gradient, and each weight or bias block.  Measured on
# MI355X (ring12, 4096 points): 4.7e-5 whole, 3.9e-4 for W0 (160-wide input
# rows), 1e-4 or less for every other block -- accumulation-order effects
GPU_GRAD_TOL, GPU_BLOCK_TOL = 5e-4, 2e-3


def layout():
    """[(woff, boff, k_in, n_out)] per layer and the parameter count"""
    out, off = [], 0
    for l in range(D + 1):
        kin = K0 if l == 0 else H
        nout = NO if l == D else H
        out.append((off, off + kin * nout, kin, nout))
        off += kin * nout + nout
    return out, off


def blocks():
    """[(name, (lo, hi))]: the flat-vector range of every weight block W_l
    and bias block b_l"""
    lay, _ = layout()
    return [(f"W{l}", (w, b)) for l, (w, b, kin, nout) in enumerate(lay)] + \
           [(f"b{l}", (b, b + nout)) for l, (w, b, kin, nout) in enumerate(lay)]


def bf16(x):
    """round to bfloat16 (nearest even) and return as float64"""
    f = np.asarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def unpack(params):
    lay, n = layout()
    assert params.size == n
    p = np.asarray(params, dtype=np.float64)
    return [(p[w:b].reshape(nout, kin), p[b:b + nout]) for w, b, kin, nout in lay]


def features(r, freqs, center_scale, bf=False):
    c = np.asarray(center_scale[:3], dtype=np.float64)
    u = (np.asarray(r, dtype=np.float64) - c) / center_scale[3]
    t = 2.0 * np.pi * u @ np.asarray(freqs, dtype=np.float64).reshape(F, 3).T
    x = np.zeros((len(u), K0))
    x[:, :F] = np.sin(t)
    x[:, F:2 * F] = np.cos(t)
    x[:, 2 * F:2 * F + 3] = u
    return bf16(x) if bf else x


def forward(params, freqs, center_scale, r, bf=False):
    """returns (y[P,16], cache) with cache = (xs, zs) for the backward pass"""
    layers = unpack(params)
    rnd = bf16 if bf else (lambda v: v)
    x = features(r, freqs, center_scale, bf)
    xs, zs = [x], []
    for l, (W, b) in enumerate(layers):
        z = x @ rnd(W).T + b
        zs.append(z)
        if l < D:
            x = rnd(silu(z))
            xs.append(x)
    return zs[-1], (xs, zs)


def amp(y):
    return np.exp(np.minimum(y[:, NOUT], 60.0))


def predict(params, freqs, center_scale, out_scale, r, bf=False):
    y, _ = forward(params, freqs, center_scale, r, bf)
    return y[:, :NOUT] * amp(y)[:, None] * np.asarray(out_scale, dtype=np.float64)


def loss_grad(params, freqs, center_scale, out_scale, r, target, eps, bf=False, plant=None):
    """(loss, flat gradient) for one batch.  plant = (kind, layer) puts a
    known mistake into one layer's backward pass, for tests that show a
    bound catches it: "no_dsilu" drops SiLU' where the gradient leaves the
    layer's input, "no_bias" leaves its bias gradient at zero"""
    layers = unpack(params)
    rnd = bf16 if bf else (lambda v: v)
    y, (xs, zs) = forward(params, freqs, center_scale, r, bf)
    P = len(y)
    t = np.asarray(target, dtype=np.float64).reshape(P, 9) / np.asarray(out_scale, dtype=np.float64)
    n = np.sqrt((t * t).sum(axis=1) + eps * eps)
    d = y[:, :NOUT] - t / n[:, None]
    dz = y[:, NOUT] - np.log(n)
    loss = float(((d * d).sum(axis=1) + dz * dz).mean())
    dy = np.zeros((P, NO))
    dy[:, :NOUT] = 2.0 * d / P
    dy[:, NOUT] = 2.0 * dz / P
    lay, n = layout()
    grad = np.zeros(n)
    g = dy
    for l in range(D, -1, -1):
        woff, boff, kin, nout = lay[l]
        W, _ = layers[l]
        grad[woff:boff] = (rnd(g).T @ xs[l]).ravel()
        grad[boff:boff + nout] = 0.0 if plant == ("no_bias", l) else g.sum(axis=0)
        if l > 0:
            dh = rnd(g) @ rnd(W)
            g = dh if plant == ("no_dsilu", l) else dh * dsilu(zs[l - 1])
    return loss, grad
