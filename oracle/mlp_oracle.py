"""CPU restatement of the learned surrogate's network (TEST INFRASTRUCTURE ONLY).

Only tests/ and bench.py's checks import this module; the product never does.

The surrogate has no counterpart in the reference, so this module restates
the network the library trains: Fourier features of a configuration
(b1, b2, exx, eyy, exy, theta), a multilayer perceptron with tanh-form GELU
hidden layers, the mean-squared-error loss and its gradient.  It runs in
float64 and, with `bf16=True`, rounds to bfloat16 (round to nearest even) at
the points where the GPU stores bfloat16: the inputs, the weights, every
hidden activation, the output gradient and every propagated gradient.

Pins (tests/test_oracle_pins.py): the float64 gradient against central
finite differences of the float64 loss; GELU and its derivative against
closed forms (gelu(0) = 0, gelu'(0) = 1/2, gelu(z) -> z and -> 0 in the
tails, gelu' against a finite difference); the bf16 rounding against exact
representable neighbours.
"""
from __future__ import annotations

import numpy as np

HARMONICS = [(1, 0), (0, 1), (1, 1), (1, -1), (2, 0), (0, 2), (2, 1), (1, 2),
             (2, -1), (1, -2), (2, 2), (2, -2), (3, 0), (0, 3), (3, 1), (1, 3)]
C_GELU = np.sqrt(2.0 / np.pi)


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out.astype(np.float64)


def f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def features(cfg):
    """[n][64] features of configurations [n][6] (see leo2d_mlp_features)."""
    cfg = np.atleast_2d(np.asarray(cfg, dtype=np.float64))
    b1, b2 = cfg[:, 0], cfg[:, 1]
    cs = np.stack([np.cos(2 * np.pi * (p * b1 + q * b2)) for p, q in HARMONICS], axis=1)
    sn = np.stack([np.sin(2 * np.pi * (p * b1 + q * b2)) for p, q in HARMONICS], axis=1)
    cols = []
    for h in range(16):
        cols += [cs[:, h], sn[:, h]]
    e = cfg[:, 2:5] / 0.003
    cols += [e[:, 0], e[:, 1], e[:, 2]]
    for a in range(3):
        for h in range(3):
            cols += [e[:, a] * cs[:, h], e[:, a] * sn[:, h]]
    t = (cfg[:, 5] - 1.9) / 1.1
    cols += [t, t * t, 1.0 / cfg[:, 5]]
    for h in range(3):
        cols += [t * cs[:, h], t * sn[:, h]]
    cols += [np.ones_like(t), t * (e[:, 0] + e[:, 1])]
    return np.stack(cols, axis=1).astype(np.float32)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3 * 0.044715 * z * z)


def unpack(params, dims):
    """Flat parameters -> [(W_l [d_{l+1}][d_l], b_l [d_{l+1}])]."""
    out, off = [], 0
    for l in range(len(dims) - 1):
        nw = dims[l + 1] * dims[l]
        W = np.asarray(params[off:off + nw], dtype=np.float64).reshape(dims[l + 1], dims[l])
        off += nw
        b = np.asarray(params[off:off + dims[l + 1]], dtype=np.float64)
        off += dims[l + 1]
        out.append((W, b))
    return out


def forward(params, dims, X, bf16_points=True):
    """Outputs Y and the cached (Hin, Z) of every layer."""
    rnd = bf16 if bf16_points else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(params, dims)
    h = rnd(X)
    cache = []
    for l, (W, b) in enumerate(layers):
        z = h @ rnd(W).T + b
        if bf16_points:
            z = f32(z)
        cache.append((h, z))
        if l + 1 < len(layers):
            h = rnd(gelu(z))
    return z, cache


def loss_and_grad(params, dims, X, T, bf16_points=True):
    """Mean squared error over all outputs and its flat gradient."""
    rnd = bf16 if bf16_points else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(params, dims)
    Y, cache = forward(params, dims, X, bf16_points)
    n, o = Y.shape
    d = Y - np.asarray(T, dtype=np.float64)
    loss = float(np.sum(d * d) / (n * o))
    g = rnd((2.0 / (n * o)) * d)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        hin, _ = cache[l]
        dW = g.T @ hin
        db = g.sum(axis=0)
        grads[l] = (dW, db)
        if l > 0:
            W = rnd(layers[l][0])
            g = rnd((g @ W) * gelu_grad(cache[l - 1][1]))
    flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
    return loss, flat
