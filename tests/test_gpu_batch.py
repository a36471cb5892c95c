"""Batched jobs (many stacked Hamiltonians in one recursion) against the oracle.

Each job is one (cell, k-point) Hamiltonian with its own R random-phase vectors
(global vector index p*R + r).  Bars: FP64 moments |dmu| <= 1e-12 mu_0 per job
(summation order only); FP32 recursion |dmu| <= 2e-5 mu_0; graph replays equal
the eager launches bit for bit (same kernels, same order).
"""
import numpy as np
import pytest

import leo2d_amd as L
from leo2d_amd import workloads as W

pytestmark = pytest.mark.gpu

CELLS = (2, 3, 5)          # (i+1, i) cells: 76, 148, 364 orbitals
KFRAC = ((0.31, 0.12), (0.77, 0.58))
R, M, SEED = 16, 64, 7


@pytest.fixture(scope="module")
def batch(oracle_mod):
    dev = L.Device(0)
    jobs = []
    for i in CELLS:
        geom = L.Geometry(L.Deck(W.tblg_cell_deck(i + 1, i, M=M)))
        assert geom.n == W.cell_orbitals(i)
        dev.set_geometry(geom)
        dev.pairs_build()
        for kf in KFRAC:
            W.assemble_rescaled(dev, geom, kf)
            rp, ci, v = dev.csr_copy()
            p = dev.batch_append()
            vecs = oracle_mod.random_phase(geom.n, R, SEED, r0=p * R)
            omu = oracle_mod.kpm(rp, ci, v, M, vecs=vecs, doubling_all=True, nthreads=8)
            jobs.append((geom.n, omu))
    yield dev, jobs
    dev.close()


@pytest.mark.parametrize("mode", [L.BATCH_ROW_GROUPS, L.BATCH_ENTRY_SLOTS, L.BATCH_ROW_CHUNKS])
@pytest.mark.parametrize("tile_rows", [64, 100, 256])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_fp64_matches_oracle(batch, mode, tile_rows, graph):
    dev, jobs = batch
    print("case mode", mode, "tile_rows", tile_rows, "graph", graph, flush=True)
    mu = dev.batch_kpm(M, R, seed=SEED, mode=mode, tile_rows=tile_rows, graph=graph)
    assert mu.shape == (len(jobs), R, M)
    for p, (n, omu) in enumerate(jobs):
        assert np.max(np.abs(mu[p] - omu)) <= 1e-12 * n, (p, mode, tile_rows)
    t = dev.batch_timing()
    assert t.jobs == len(jobs) and t.rows == sum(n for n, _ in jobs)
    assert t.mode == mode and t.tile_rows == tile_rows and t.graph == int(graph)


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("cw", ["1", "0"])
def test_batch_sell_matches_oracle(batch, graph, cw, monkeypatch):
    """SELL-4-sigma tiles cut at the job boundaries (six jobs of 76 to 364
    orbitals), partial sums per tile, staging copies by a wavefront of their
    own (cw = 1, tiles of at most 56 rows) or by every wavefront (cw = 0):
    FP64 moments within 1e-12 mu_0 of the oracle for every job, eager and
    captured.  The FP32 recursion at 16 vectors per job (128-byte row chunks)
    runs only with the copy wavefront: within 2e-5 mu_0 there, an argument
    error without it."""
    monkeypatch.setenv("LEO2D_SELL_CW", cw)
    dev, jobs = batch
    print("case SELL graph", graph, "copy wave", cw, flush=True)
    mu = dev.batch_kpm(M, R, seed=SEED, mode=L.BATCH_SELL, graph=graph)
    for p, (n, omu) in enumerate(jobs):
        assert np.max(np.abs(mu[p] - omu)) <= 1e-12 * n, (p, graph)
    t = dev.batch_timing()
    assert t.mode == L.BATCH_SELL and 0 < t.tile_rows <= (56 if cw == "1" else 64) and t.ntiles >= len(jobs)
    if cw == "1":
        mu32 = dev.batch_kpm(M, R, seed=SEED, mode=L.BATCH_SELL, precision=L.PREC_MIXED, graph=graph)
        for p, (n, omu) in enumerate(jobs):
            assert np.max(np.abs(mu32[p] - omu)) <= 2e-5 * n, (p, graph)
    else:
        with pytest.raises(L.Leo2dError) as e:
            dev.batch_kpm(M, R, seed=SEED, mode=L.BATCH_SELL, precision=L.PREC_MIXED)
        assert e.value.code == L.ERR_ARG
    if graph:
        eager = dev.batch_kpm(M, R, seed=SEED, mode=L.BATCH_SELL, graph=False)
        for rep in range(3):
            g = dev.batch_kpm(M, R, seed=SEED, mode=L.BATCH_SELL, graph=True)
            assert np.array_equal(g, eager), rep
            dev.kpm(16, random_phase=(4, rep), scheme=L.SCHEME_DOUBLING)


@pytest.mark.parametrize("Rv", [4, 8, 32])
def test_batch_other_block_widths(oracle_mod, Rv):
    """R = 4, 8, 32 vectors per job on two jobs of different size."""
    dev = L.Device(0)
    try:
        want = []
        for i in (2, 4):
            geom = L.Geometry(L.Deck(W.tblg_cell_deck(i + 1, i, M=32)))
            dev.set_geometry(geom)
            dev.pairs_build()
            W.assemble_rescaled(dev, geom, (0.2, 0.45))
            rp, ci, v = dev.csr_copy()
            p = dev.batch_append()
            vecs = oracle_mod.random_phase(geom.n, Rv, 11, r0=p * Rv)
            want.append((geom.n, oracle_mod.kpm(rp, ci, v, 32, vecs=vecs, doubling_all=True, nthreads=8)))
        for mode in (L.BATCH_ROW_GROUPS, L.BATCH_ENTRY_SLOTS, L.BATCH_ROW_CHUNKS):
            if mode == L.BATCH_ROW_CHUNKS and Rv < 16:
                # 16-lane row slots need 16 or 32 vectors per job
                with pytest.raises(L.Leo2dError) as e:
                    dev.batch_kpm(32, Rv, seed=11, mode=mode, graph=False)
                assert e.value.code == L.ERR_ARG
                continue
            mu = dev.batch_kpm(32, Rv, seed=11, mode=mode, graph=False)
            for p, (n, omu) in enumerate(want):
                assert np.max(np.abs(mu[p] - omu)) <= 1e-12 * n, (Rv, mode, p)
        # SELL tiles: 16 or 32 vectors per job; the FP32 recursion at 32
        if Rv < 16:
            with pytest.raises(L.Leo2dError) as e:
                dev.batch_kpm(32, Rv, seed=11, mode=L.BATCH_SELL, graph=False)
            assert e.value.code == L.ERR_ARG
        else:
            for prec, tol in ((L.PREC_FP64, 1e-12), (L.PREC_MIXED, 2e-5)):
                mu = dev.batch_kpm(32, Rv, seed=11, mode=L.BATCH_SELL, precision=prec, graph=False)
                for p, (n, omu) in enumerate(want):
                    assert np.max(np.abs(mu[p] - omu)) <= tol * n, (Rv, prec, p)
    finally:
        dev.close()


@pytest.mark.parametrize("mode", [L.BATCH_ROW_GROUPS, L.BATCH_ROW_CHUNKS])
def test_batch_mixed_precision(batch, mode):
    dev, jobs = batch
    mu = dev.batch_kpm(M, R, seed=SEED, precision=L.PREC_MIXED, graph=True, mode=mode)
    for p, (n, omu) in enumerate(jobs):
        assert np.max(np.abs(mu[p] - omu)) <= 2e-5 * n, (p, mode)


def test_batch_autotune_and_cache(batch):
    dev, jobs = batch
    mu = dev.batch_kpm(M, R, seed=SEED, tune=True)
    t = dev.batch_timing()
    assert t.candidates > 0 and t.tune_ms > 0
    for p, (n, omu) in enumerate(jobs):
        assert np.max(np.abs(mu[p] - omu)) <= 1e-12 * n
    mu2 = dev.batch_kpm(M, R, seed=SEED, tune=True)
    t2 = dev.batch_timing()
    assert t2.candidates == 0 and (t2.mode, t2.tile_rows, t2.grid) == (t.mode, t.tile_rows, t.grid)
    assert np.array_equal(mu, mu2)


def test_batch_graph_replays_match_eager(batch):
    """The captured recursion is replayed several times with other device work
    in between; every replay equals the eager launches bit for bit."""
    dev, jobs = batch
    eager = dev.batch_kpm(M, R, seed=SEED, mode=L.BATCH_ENTRY_SLOTS, tile_rows=128, graph=False)
    for rep in range(4):
        g = dev.batch_kpm(M, R, seed=SEED, mode=L.BATCH_ENTRY_SLOTS, tile_rows=128, graph=True)
        assert np.array_equal(g, eager), rep
        dev.kpm(16, random_phase=(4, rep), scheme=L.SCHEME_DOUBLING)   # unrelated eager work on the same stream


def test_batch_append_after_run_and_clear(oracle_mod):
    dev = L.Device(0)
    try:
        geom = L.Geometry(L.Deck(W.tblg_cell_deck(4, 3, M=32)))
        dev.set_geometry(geom)
        dev.pairs_build()
        W.assemble_rescaled(dev, geom, (0.1, 0.9))
        rp, ci, v = dev.csr_copy()
        assert dev.batch_append() == 0
        dev.batch_kpm(32, 8, seed=2, graph=True)
        assert dev.batch_append() == 1                  # the same H again, stacked after the first run
        mu = dev.batch_kpm(32, 8, seed=2, graph=True)
        for p in range(2):
            omu = oracle_mod.kpm(rp, ci, v, 32, vecs=oracle_mod.random_phase(geom.n, 8, 2, r0=p * 8),
                                 doubling_all=True, nthreads=8)
            assert np.max(np.abs(mu[p] - omu)) <= 1e-12 * geom.n
        dev.batch_clear()
        assert dev.batch_info() == (0, 0, 0)
        with pytest.raises(L.Leo2dError):
            dev.batch_kpm(32, 8)
    finally:
        dev.close()


def test_batch_argument_errors(batch):
    dev, _ = batch
    for kw in ({"R": 3}, {"R": 64}, {"M": 30}, {"grid": 12}, {"tile_rows": 8}):
        args = {"M": M, "R": R}
        args.update(kw)
        with pytest.raises(L.Leo2dError):
            dev.batch_kpm(args.pop("M"), args.pop("R"), **args)


def test_batch_target_starts(batch, oracle_mod):
    """Local-DoS starts: unit vectors on R orbitals of each job."""
    dev, jobs = batch
    rng = np.random.default_rng(3)
    Rt = 4
    targets = np.array([rng.choice(n, size=Rt, replace=False) for n, _ in jobs], dtype=np.int32)
    mu = dev.batch_kpm(M, Rt, targets=targets, graph=True)
    # the oracle needs each job's matrix: rebuild them in the fixture's order
    p = 0
    for i in CELLS:
        geom = L.Geometry(L.Deck(W.tblg_cell_deck(i + 1, i, M=M)))
        d2 = L.Device(0)
        try:
            d2.set_geometry(geom)
            d2.pairs_build()
            for kf in KFRAC:
                W.assemble_rescaled(d2, geom, kf)
                rp, ci, v = d2.csr_copy()
                omu = oracle_mod.kpm(rp, ci, v, M, targets=targets[p], doubling_all=True, nthreads=8)
                assert np.max(np.abs(mu[p] - omu)) <= 1e-12, p
                assert np.allclose(mu[p][:, 0], 1.0)
                p += 1
        finally:
            d2.close()
    with pytest.raises(L.Leo2dError):
        bad = targets.copy()
        bad[0, 0] = jobs[0][0]          # one past the first job's last orbital
        dev.batch_kpm(M, Rt, targets=bad)


def _gamma_job(dev, geom):
    """Gamma-point job: real H (no Bloch phase), Gershgorin interval with 1 % padding."""
    job = geom.jobs()[0]
    job.complex_matrix = 0
    job.k[0], job.k[1] = 0.0, 0.0
    job.energy_rescale, job.energy_shift = 1.0, 0.0
    dev.assemble(job, L.PATTERN_PRUNED)
    lo, hi = dev.spectral_bounds()
    job.energy_rescale, job.energy_shift = 1.01 * (hi - lo) / 2.0, -(hi + lo) / 2.0
    dev.assemble(job, L.PATTERN_PRUNED)


@pytest.mark.parametrize("Rt", [4, 16, 32])
def test_batch_real_target_starts(oracle_mod, Rt):
    """Real H (Gamma point) with unit-vector starts runs on real vector blocks
    (FP64 and FP32 elements): every lane layout, both precisions, and the SELL
    tiles with two FP64 vectors per lane at 32 vectors per job, against the
    oracle's faithful KPM from the same targets (FP64 within 1e-12, FP32
    recursion within 2e-5 of mu_0 = 1)."""
    dev = L.Device(0)
    try:
        rng = np.random.default_rng(Rt)
        want, tg = [], []
        for i in (3, 5):
            geom = L.Geometry(L.Deck(W.tblg_cell_deck(i + 1, i, M=M)))
            dev.set_geometry(geom)
            dev.pairs_build()
            _gamma_job(dev, geom)
            rp, ci, v = dev.csr_copy()
            assert not np.iscomplexobj(v)
            t = rng.choice(geom.n, size=Rt, replace=False).astype(np.int32)
            tg.append(t)
            want.append(oracle_mod.kpm(rp, ci, v, M, targets=t, doubling_all=True, nthreads=8))
            dev.batch_append()
        tg = np.array(tg)
        for mode in (L.BATCH_ROW_GROUPS, L.BATCH_ENTRY_SLOTS, L.BATCH_ROW_CHUNKS, L.BATCH_SELL):
            if mode == L.BATCH_ROW_CHUNKS and Rt < 16:
                continue
            for prec, tol in ((L.PREC_FP64, 1e-12), (L.PREC_MIXED, 2e-5)):
                if mode == L.BATCH_SELL and (Rt != 32 or prec != L.PREC_FP64):
                    continue   # real SELL tiles: FP64 vectors, 32 per job
                print("case R", Rt, "mode", mode, "prec", prec, flush=True)
                mu = dev.batch_kpm(M, Rt, targets=tg, precision=prec, mode=mode, graph=False)
                for p, omu in enumerate(want):
                    assert np.max(np.abs(mu[p] - omu)) <= tol, (Rt, mode, prec, p)
    finally:
        dev.close()


def test_batch_tuning_cache_per_vector_type():
    """On a real-H stack, random-phase starts run complex vector blocks and
    target starts real ones: a shape tuned for one element type is not reused
    for the other (the cache is keyed on it), a repeat of either is, and
    restacking the batch drops the cached shapes."""
    d2 = L.Device(0)
    try:
        geom = L.Geometry(L.Deck(W.tblg_cell_deck(4, 3, M=M)))
        d2.set_geometry(geom)
        d2.pairs_build()
        _gamma_job(d2, geom)
        d2.batch_append()
        tg = np.array([[0, 5, 9, 11, 20, 31, 40, 41, 50, 60, 70, 80, 90, 100, 110, 120]], dtype=np.int32)
        d2.batch_kpm(M, 16, seed=1, tune=True)
        assert d2.batch_timing().candidates > 0
        d2.batch_kpm(M, 16, targets=tg, tune=True)
        assert d2.batch_timing().candidates > 0          # real blocks: tuned afresh
        d2.batch_kpm(M, 16, targets=tg, tune=True)
        assert d2.batch_timing().candidates == 0         # cached
        d2.batch_kpm(M, 16, seed=1, tune=True)
        assert d2.batch_timing().candidates == 0         # cached
        d2.batch_append()                                # restack: cached shapes dropped
        d2.batch_kpm(M, 16, seed=1, tune=True)
        assert d2.batch_timing().candidates > 0
    finally:
        d2.close()


def test_batch_c2_cells(oracle_mod):
    """Cells from the batched workload's range (i = 9 and 13: 1,084 and 2,188
    orbitals), four Bloch k-points each, 16 random-phase vectors per job,
    M = 1024, tuned launch shape: every job's FP64 moments within 1e-12 mu_0
    of the oracle."""
    dev = L.Device(0)
    try:
        rng = np.random.default_rng(1)
        want = []
        Mb, Rb = 1024, 16
        for i in (9, 13):
            geom = L.Geometry(L.Deck(W.tblg_cell_deck(i + 1, i, M=Mb)))
            assert geom.n == W.cell_orbitals(i)
            dev.set_geometry(geom)
            dev.pairs_build()
            for kf in rng.random((4, 2)):
                W.assemble_rescaled(dev, geom, kf)
                rp, ci, v = dev.csr_copy()
                p = dev.batch_append()
                vecs = oracle_mod.random_phase(geom.n, Rb, 5, r0=p * Rb)
                want.append((geom.n, oracle_mod.kpm(rp, ci, v, Mb, vecs=vecs, doubling_all=True, nthreads=16)))
        mu = dev.batch_kpm(Mb, Rb, seed=5, tune=True)
        for p, (n, omu) in enumerate(want):
            assert np.max(np.abs(mu[p] - omu)) <= 1e-12 * n, p
        # SELL tiles, several per job, walked in both directions over the 512 steps
        mu = dev.batch_kpm(Mb, Rb, seed=5, mode=L.BATCH_SELL, graph=True)
        assert dev.batch_timing().ntiles > len(want)
        for p, (n, omu) in enumerate(want):
            assert np.max(np.abs(mu[p] - omu)) <= 1e-12 * n, p
    finally:
        dev.close()
