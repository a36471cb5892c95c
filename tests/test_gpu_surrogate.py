"""What the trained surrogate gains over exact local KPM alone.

Small instance of the learned-surrogate workload: M = 64 local moments of
orbitals of (i+1, i) commensurate cells with twist in [2.0, 3.0] deg, random
interlayer shift and strain.  Every target orbital of a cell is a sample (its
stacking at its unit cell and its sublattice block).  Exact moments come from
the library's batched KPM (target starts); an ensemble of four 3 x 256 GELU
networks is trained on the GPU and evaluated on held-out samples.

Checks: the exact generator returns local-DoS moments (mu_0 = 1, |mu_n| <= 1)
that agree with the single-job KPM path; a sample's moments equal those of its
configuration evaluated afresh at an origin-cell orbital (same stacking field,
same environment) -- this pins the stacking-field convention; training lowers
the loss; the ensemble beats the best configuration-blind predictor (the
per-block training mean) by a wide margin; its spread flags the samples it gets
most wrong; the fallback replaces exactly the flagged ones by exact moments;
and the per-sample times of both (exact and ensemble) are recorded as test
properties.
"""
import time

import numpy as np
import pytest

import leo2d_amd as L
from leo2d_amd import surrogate as S

pytestmark = pytest.mark.gpu

M = 64
N_JOBS, R = 64, 32
N_TEST = 256


@pytest.fixture(scope="module")
def data():
    dev = L.Device(0)
    rng = np.random.default_rng(21)
    ex = S.ExactLocalKPM(dev, M)
    t0 = time.perf_counter()
    d = ex.samples(N_JOBS, rng, R=R, theta_range=(2.0, 3.0))
    t_exact = time.perf_counter() - t0
    yield dev, ex, d, t_exact
    dev.close()


def test_exact_local_moments(data):
    dev, ex, d, _ = data
    mom = d["moments"]
    assert mom.shape == (N_JOBS * R, M)
    assert np.allclose(mom[:, 0], 1.0)
    assert np.max(np.abs(mom)) <= 1.0 + 1e-9
    # one sample through the single-job path
    q = 7
    g = ex._geom(int(d["cell"][q]))
    dev.set_geometry(g)
    dev.pairs_build()
    job = g.jobs()[0]
    job.shift[1][0], job.shift[1][1] = d["shift"][q]
    c = d["cfg"][q]
    job.uniform_strain = 1
    job.u[0][0], job.u[0][1], job.u[1][0], job.u[1][1] = c[2], c[4], c[4], c[3]
    job.complex_matrix = 0
    job.energy_rescale, job.energy_shift = S.RESCALE, 0.0
    dev.assemble(job, L.PATTERN_PRUNED)
    mu = dev.kpm(M, targets=[int(d["target"][q])], scheme=L.SCHEME_DOUBLING)
    assert np.max(np.abs(mu[0] - mom[q])) <= 1e-12
    # the same configurations evaluated at origin-cell orbitals, two per block
    pick = np.concatenate([np.nonzero(d["block"] == k)[0][:2] for k in range(4)])
    again = ex.moments_at(d["cfg"][pick], d["block"][pick])
    assert np.max(np.abs(again - mom[pick])) <= 1e-9, np.max(np.abs(again - mom[pick]), axis=1)


def test_surrogate_gain(data, record_property):
    dev, ex, d, t_exact = data
    x = S.site_features(d["cfg"], d["block"])
    t = d["moments"].astype(np.float32)
    n_tr = len(x) - N_TEST
    xtr, ttr, xte, tte = x[:n_tr], t[:n_tr], x[n_tr:], t[n_tr:]
    btr, bte = d["block"][:n_tr], d["block"][n_tr:]
    nets, losses = S.train_ensemble(xtr, ttr, members=4, hidden=256, layers=3, epochs=80, batch=128, lr=1e-3)
    try:
        for ls in losses:
            assert ls[-1] < 0.2 * ls[0], ls[[0, -1]]
        t0 = time.perf_counter()
        mean, spread = L.mlp_ensemble(nets, xte)
        t_mlp = time.perf_counter() - t0
        err = np.sqrt(np.mean((mean - tte) ** 2, axis=1))           # per sample
        block_mean = np.stack([ttr[btr == k].mean(axis=0) for k in range(4)])
        base = np.sqrt(np.mean((block_mean[bte] - tte) ** 2, axis=1))
        print(f"rms error: ensemble {err.mean():.3e}, per-block mean predictor {base.mean():.3e}", flush=True)
        assert err.mean() < 0.35 * base.mean()
        # the spread flags the worst predictions
        top = spread >= np.quantile(spread, 0.9)
        assert err[top].mean() > err[~top].mean()
        # fallback: exactly the flagged samples become exact
        thr = float(np.quantile(spread, 0.9))
        cfg_te = d["cfg"][n_tr:]
        out, fb, sp, rep = S.predict_with_fallback(nets, ex, cfg_te, bte, threshold=thr)
        assert np.array_equal(fb, sp > thr) and rep["fallback"] == int(fb.sum()) > 0
        assert np.max(np.abs(out[fb] - tte[fb])) <= 1e-6
        final = np.sqrt(np.mean((out - tte) ** 2, axis=1))
        assert final.max() <= err.max() + 1e-6
        # speed: exact local KPM per sample vs ensemble inference per sample, recorded
        # (a wall-clock ratio depends on the machine's load, so it is not asserted)
        per_exact = t_exact / len(x)
        per_mlp = t_mlp / N_TEST
        print(f"per sample: exact {per_exact * 1e3:.3f} ms, ensemble {per_mlp * 1e3:.4f} ms "
              f"({per_exact / per_mlp:.0f}x)", flush=True)
        record_property("exact_ms_per_sample", per_exact * 1e3)
        record_property("ensemble_ms_per_sample", per_mlp * 1e3)
        record_property("speed_ratio", per_exact / per_mlp)
    finally:
        for m in nets:
            m.close()
