"""The surrogate's BF16 MFMA kernels against the CPU oracle network.

Bars: the oracle emulates every bfloat16 rounding point of the GPU, so what
is left is FP32 (GPU) vs FP64 (oracle) accumulation and the occasional
activation that rounds to the neighbouring bfloat16 value.  Forward outputs:
relative Frobenius error <= 5e-3; gradients: relative Frobenius error per
layer <= 2e-2 (bias gradients sum bf16 values over the batch).  A transposed
operand, a wrong lane layout or a dropped GELU' term gives errors of order 1.
"""
import numpy as np
import pytest

import leo2d_amd as L
from oracle import mlp_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def _data(rng, n, d_in, d_out):
    cfg = np.column_stack([rng.random(n), rng.random(n), rng.normal(0, 0.003, (n, 3)), rng.uniform(0.8, 3, n)])
    x = L.mlp_features(cfg)[:, :d_in] if d_in <= 64 else rng.normal(size=(n, d_in)).astype(np.float32)
    t = rng.normal(0, 0.3, (n, d_out)).astype(np.float32)
    return x, t


def _layer_slices(dims):
    out, off = [], 0
    for l in range(len(dims) - 1):
        nw = dims[l + 1] * dims[l]
        out.append((slice(off, off + nw), slice(off + nw, off + nw + dims[l + 1])))
        off += nw + dims[l + 1]
    return out


@pytest.mark.parametrize("shape", [(64, 96, 2, 40, 150), (64, 256, 3, 256, 256), (48, 64, 1, 130, 70)])
def test_mlp_forward_matches_oracle(shape):
    d_in, hidden, layers, out, n = shape
    print("case", shape, flush=True)
    rng = np.random.default_rng(10)
    net = L.Mlp(0, d_in, hidden, layers, out, seed=3)
    try:
        x, _ = _data(rng, n, d_in, out)
        y = net.forward(x)
        ref, _ = O.forward(net.params(), net.dims, x)
        assert _rel(y, ref) <= 5e-3, _rel(y, ref)
        # every row and column of the output tile grid is written (no stale tiles)
        assert np.all(np.isfinite(y))
    finally:
        net.close()


@pytest.mark.parametrize("shape", [(64, 96, 2, 40, 150), (64, 256, 3, 512, 128)])
def test_mlp_gradients_match_oracle(shape):
    d_in, hidden, layers, out, n = shape
    print("case", shape, flush=True)
    rng = np.random.default_rng(11)
    net = L.Mlp(0, d_in, hidden, layers, out, seed=5)
    try:
        x, t = _data(rng, n, d_in, out)
        loss, g = net.gradients(x, t)
        oloss, og = O.loss_and_grad(net.params(), net.dims, x, t)
        assert abs(loss - oloss) <= 1e-3 * oloss
        for l, (sw, sb) in enumerate(_layer_slices(net.dims)):
            assert _rel(g[sw], og[sw]) <= 2e-2, (l, "W", _rel(g[sw], og[sw]))
            assert _rel(g[sb], og[sb]) <= 2e-2, (l, "b", _rel(g[sb], og[sb]))
    finally:
        net.close()


def test_mlp_set_params_round_trip_and_training_descends():
    rng = np.random.default_rng(12)
    net = L.Mlp(0, 64, 128, 2, 32, seed=9)
    try:
        p = net.params()
        net.set_params(p * 0.5)
        assert np.array_equal(net.params(), (p * 0.5).astype(np.float32))
        # a smooth target of the features: the loss must fall by an order of magnitude
        cfg = np.column_stack([rng.random(2048), rng.random(2048), rng.normal(0, 0.003, (2048, 3)),
                               rng.uniform(0.8, 3, 2048)])
        x = L.mlp_features(cfg)
        A = rng.normal(0, 0.2, (64, 32)).astype(np.float32)
        t = np.tanh(x @ A).astype(np.float32)
        losses = net.train(x, t, epochs=40, batch=256, lr=2e-3, seed=1)
        assert np.all(np.isfinite(losses))
        assert losses[-1] < 0.1 * losses[0], losses
    finally:
        net.close()


def test_mlp_ensemble_mean_and_spread():
    rng = np.random.default_rng(13)
    nets = [L.Mlp(0, 64, 64, 2, 48, seed=s) for s in range(4)]
    try:
        x, _ = _data(rng, 300, 64, 48)
        ys = np.stack([m.forward(x) for m in nets])
        mean, spread = L.mlp_ensemble(nets, x)
        assert np.max(np.abs(mean - ys.mean(axis=0))) <= 1e-5 * max(1.0, np.max(np.abs(ys)))
        want = np.sqrt(np.mean(ys.var(axis=0), axis=1))
        assert np.max(np.abs(spread - want)) <= 1e-5 * max(1.0, np.max(want))
        _, s2 = L.mlp_ensemble(nets, x, with_mean=False)
        assert np.array_equal(s2, spread)
        _, s4 = L.mlp_ensemble(nets, x, blocks=4)
        want4 = np.sqrt(np.mean(ys.var(axis=0).reshape(len(x), 4, 12), axis=2))
        assert s4.shape == (len(x), 4) and np.max(np.abs(s4 - want4)) <= 1e-5 * max(1.0, np.max(want4))
    finally:
        for m in nets:
            m.close()


def test_train_ensemble_matches_sequential_training():
    """Three networks trained together (one captured graph per network and epoch,
    replayed on each network's stream, the Adam step from a device counter) end
    with the parameters and per-epoch losses of the same networks trained one
    after another by leo2d_mlp_train with the same seeds, bit for bit; a fourth
    epoch count exercises the partial last mini-batch (1000 rows, batches of 96)."""
    rng = np.random.default_rng(17)
    x, t = _data(rng, 1000, 64, 40)
    seeds = [101, 202, 303]
    together = [L.Mlp(0, 64, 96, 2, 40, seed=s) for s in (1, 2, 3)]
    alone = [L.Mlp(0, 64, 96, 2, 40, seed=s) for s in (1, 2, 3)]
    try:
        for rep in range(2):
            lt = L.mlp_train_ensemble(together, x, t, seeds, epochs=3, batch=96, lr=2e-3)
            for k, net in enumerate(alone):
                la = net.train(x, t, epochs=3, batch=96, lr=2e-3, seed=seeds[k])
                assert np.array_equal(lt[k], la), (rep, k)
                assert np.array_equal(together[k].params(), net.params()), (rep, k)
            assert np.all(lt[:, -1] < lt[:, 0])
    finally:
        for net in together + alone:
            net.close()
