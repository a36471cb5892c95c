"""What the trained surrogate gains over exact local KPM alone.

Small instance of the learned-surrogate workload: M = 64 local moments of the
four origin-cell orbitals of (i+1, i) commensurate cells with twist in
[2.0, 3.0] deg, random interlayer shift and strain.  Exact moments come from
the library's batched KPM (target starts); an ensemble of four 3 x 256 GELU
networks is trained on the GPU and evaluated on held-out configurations.

Checks: the exact generator returns local-DoS moments (mu_0 = 1, |mu_n| <= 1)
and agrees with the single-job KPM path; training lowers the loss; the
ensemble beats the best configuration-blind predictor (the training mean) by
a wide margin; its spread flags the configurations it gets most wrong; the
fallback replaces exactly the flagged ones by exact moments; and the networks
answer many times faster than the exact computation they replace.
"""
import time

import numpy as np
import pytest

import leo2d_amd as L
from leo2d_amd import surrogate as S

pytestmark = pytest.mark.gpu

M = 64
N_TRAIN, N_TEST = 1536, 256


@pytest.fixture(scope="module")
def data():
    dev = L.Device(0)
    rng = np.random.default_rng(21)
    cfg = S.sample_configs(N_TRAIN + N_TEST, rng, theta_range=(2.0, 3.0))
    ex = S.ExactLocalKPM(dev, M)
    t0 = time.perf_counter()
    mom = ex.moments(cfg)
    t_exact = time.perf_counter() - t0
    yield dev, ex, cfg, mom, t_exact
    dev.close()


def test_exact_local_moments(data):
    dev, ex, cfg, mom, _ = data
    assert np.allclose(mom[:, :, 0], 1.0)
    assert np.max(np.abs(mom)) <= 1.0 + 1e-9
    # one configuration through the single-job path gives the same moments
    q = 7
    i = S.nearest_cell(cfg[q, 5])
    g = ex._geom(i)
    dev.set_geometry(g)
    dev.pairs_build()
    job = g.jobs()[0]
    job.shift[1][0], job.shift[1][1] = cfg[q, 0], cfg[q, 1]
    job.uniform_strain = 1
    job.u[0][0], job.u[0][1], job.u[1][0], job.u[1][1] = cfg[q, 2], cfg[q, 4], cfg[q, 4], cfg[q, 3]
    job.complex_matrix = 0
    job.energy_rescale, job.energy_shift = S.RESCALE, 0.0
    dev.assemble(job, L.PATTERN_PRUNED)
    mu = dev.kpm(M, targets=g._origin, scheme=L.SCHEME_DOUBLING)
    assert np.max(np.abs(mu - mom[q])) <= 1e-12


def test_surrogate_gain(data):
    dev, ex, cfg, mom, t_exact = data
    x = L.mlp_features(cfg)
    t = mom.reshape(len(cfg), -1).astype(np.float32)
    xtr, ttr, xte, tte = x[:N_TRAIN], t[:N_TRAIN], x[N_TRAIN:], t[N_TRAIN:]
    nets, losses = S.train_ensemble(xtr, ttr, members=4, hidden=256, layers=3, epochs=80, batch=128, lr=1e-3)
    try:
        for ls in losses:
            assert ls[-1] < 0.2 * ls[0], ls[[0, -1]]
        t0 = time.perf_counter()
        mean, sp4 = L.mlp_ensemble(nets, xte, blocks=4)
        t_mlp = time.perf_counter() - t0
        spread = sp4.max(axis=1)
        err = np.sqrt(np.mean((mean - tte) ** 2, axis=1))           # per configuration
        base = np.sqrt(np.mean((ttr.mean(axis=0) - tte) ** 2, axis=1))
        print(f"rms error: ensemble {err.mean():.3e}, training-mean predictor {base.mean():.3e}", flush=True)
        assert err.mean() < 0.35 * base.mean()
        # the spread flags the worst predictions
        top = spread >= np.quantile(spread, 0.9)
        assert err[top].mean() > err[~top].mean()
        # fallback: exactly the flagged configurations become exact
        thr = float(np.quantile(spread, 0.9))
        test_cfg = cfg[N_TRAIN:]
        out, fb, sp, rep = S.predict_with_fallback(nets, ex, test_cfg, threshold=thr)
        assert np.array_equal(fb, sp > thr) and rep["fallback"] == int(fb.sum()) > 0
        assert np.max(np.abs(out[fb] - mom[N_TRAIN:][fb])) <= 1e-12
        final = np.sqrt(np.mean((out.reshape(len(test_cfg), -1) - tte) ** 2, axis=1))
        assert final.max() <= err.max() + 1e-7
        # speed: exact local KPM per configuration vs ensemble inference per configuration
        per_exact = t_exact / len(cfg)
        per_mlp = t_mlp / N_TEST
        print(f"per configuration: exact {per_exact * 1e3:.3f} ms, ensemble {per_mlp * 1e3:.4f} ms "
              f"({per_exact / per_mlp:.0f}x)", flush=True)
        assert per_exact > 10 * per_mlp
    finally:
        for m in nets:
            m.close()
