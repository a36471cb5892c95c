"""Learned local-moment surrogate with exact local-KPM fallback.

A local stacking configuration c = (b1, b2, exx, eyy, exy, theta) is: the
shift of sheet 2 against sheet 1 at a unit cell, in fractional lattice units
of sheet 2; a symmetric strain tensor applied to both sheets; and the twist
in degrees.  In a rigidly twisted bilayer the shift is the linear field
b(r) = A^-1 (I - R^-1) r + s (mod 1), with r the unit cell's position, A the
lattice, R the twist and s the shift applied to sheet 2.

Exact local moments of an orbital come from a commensurate (i+1, i)
twisted-bilayer cell whose twist is nearest theta, with sheet 2 shifted by s
(setConfigPositions shifts, src/locality.cpp:3577-3660), the strain applied
as UNIFORM_STRAIN (src/locality.cpp:3849-3857), GrapheneSimple couplings and a
fixed rescale a = 10.5 eV, through the library's batched KPM with target
starts.  Every orbital of such a cell is one sample: its configuration is b
at its own unit cell, and its sublattice (sheet 1 A, B; sheet 2 A, B) is the
block index.  Two orbitals with the same configuration and block have the
same local environment (the stacking field around each is the same linear
field), so a configuration can also be evaluated on demand at an origin-cell
orbital with the matching shift; that is the fallback.

An ensemble of MLPs (64 Fourier features + a one-hot block -> 3 x 256 GELU ->
M moments; bfloat16 weights, FP32 accumulation, trained on the GPU) predicts
the moments.  A site whose ensemble spread (RMS over the M moments of the
ensemble standard deviation) exceeds a threshold falls back to the exact
computation, which also scores the networks' error there.
"""
from __future__ import annotations

import math
import time

import numpy as np

import leo2d_amd as L
from leo2d_amd import workloads as W

GR_A = 2.4768                      # graphene lattice constant of the coupling model (A)
LAT = np.array([[GR_A, 0.0], [0.5 * GR_A, 0.5 * math.sqrt(3.0) * GR_A]])   # rows a1, a2
RESCALE = 10.5                     # eV, above every Gershgorin bound of the strained cells used
N_IN = L.MLP_FEATURES + 8          # features + one-hot block (4) + zero padding to a multiple of 8
TAU_B = np.array([0.5 * GR_A, GR_A / (2.0 * math.sqrt(3.0))])


def rot(t):
    return np.array([[math.cos(t), -math.sin(t)], [math.sin(t), math.cos(t)]])


def site_features(cfg, block) -> np.ndarray:
    """[n][N_IN] network inputs: the configuration's Fourier features and the block."""
    cfg = np.atleast_2d(cfg)
    block = np.asarray(block, dtype=np.int64)
    f = np.zeros((cfg.shape[0], N_IN), dtype=np.float32)
    f[:, :L.MLP_FEATURES] = L.mlp_features(cfg)
    f[np.arange(cfg.shape[0]), L.MLP_FEATURES + block] = 1.0
    return f


def local_shifts(geom, shift=(0.0, 0.0)):
    """Stacking shift b [n][2] (mod 1) at every orbital's unit cell and the
    orbital's block 2 sheet + sublattice, for the geometry with sheet 2
    shifted by `shift` (lattice units of sheet 2).  Sheet 1 is unrotated."""
    i2g, i2p = geom.arrays()
    grid, pos = i2g.reshape(-1, 4), i2p.reshape(-1, 3)[:, :2]
    th = geom.info.angle[1]
    R2 = rot(th)
    Rinv = rot(-th)
    s = np.asarray(shift, dtype=np.float64)
    sheet2 = grid[:, 3] == 1
    off = np.where((grid[:, 2] == 1)[:, None], np.where(sheet2[:, None], (R2 @ TAU_B)[None, :], TAU_B[None, :]), 0.0)
    r = pos - off
    r = r + np.where(sheet2[:, None], (R2 @ (LAT.T @ s))[None, :], 0.0)   # sheet 2 cells moved by the shift
    d = r - r @ Rinv.T
    b = (np.linalg.solve(LAT.T, d.T).T + s) % 1.0
    return b, (2 * grid[:, 3] + grid[:, 2]).astype(np.int32)


def cell_angle_deg(i: int) -> float:
    """Twist of the (i+1, i) commensurate cell: cos theta = (N^2 + 4NM + M^2) / (2 (N^2 + NM + M^2))."""
    m, n = i + 1, i
    return math.degrees(math.acos((n * n + 4 * n * m + m * m) / (2.0 * (n * n + n * m + m * m))))


def nearest_cell(theta_deg: float, imin: int = 2, imax: int = 200) -> int:
    """Index i of the (i+1, i) cell whose twist is nearest theta (angles fall monotonically with i)."""
    lo, hi = imin, imax
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if cell_angle_deg(mid) > theta_deg:
            lo = mid
        else:
            hi = mid
    return lo if abs(cell_angle_deg(lo) - theta_deg) <= abs(cell_angle_deg(hi) - theta_deg) else hi


def sample_configs(n: int, rng, theta_range=(0.8, 3.0), strain_sd=0.003, snap=True, log_theta=False) -> np.ndarray:
    """Uniform shift, N(0, strain_sd) symmetric strain, twist uniform (or
    log-uniform) in theta_range, snapped to the nearest commensurate cell's
    twist when snap."""
    b = rng.random((n, 2))
    e = rng.normal(0.0, strain_sd, (n, 3))
    if log_theta:
        th = np.exp(rng.uniform(math.log(theta_range[0]), math.log(theta_range[1]), n))
    else:
        th = rng.uniform(theta_range[0], theta_range[1], n)
    if snap:
        th = np.array([cell_angle_deg(nearest_cell(t)) for t in th])
    return np.column_stack([b, e, th])


class ExactLocalKPM:
    """Exact local moments on one device handle: random samples (every
    target orbital of a cell is a sample) and on-demand configurations."""

    def __init__(self, dev: L.Device, M: int, chunk: int = 256, precision=L.PREC_FP64, mode=None, progress=None):
        self.dev, self.M, self.chunk, self.precision = dev, M, chunk, precision
        self.mode = mode           # batched layout (None: the default of batch_mode below)
        self.progress = progress   # called as progress(jobs done, jobs) after every batch
        self._geoms = {}

    def batch_mode(self, R, rows=0):
        """Default batched layout: SELL-4-sigma tiles cut at job boundaries for 32
        FP64 vectors per job (c4's 256 cells: 2.5 s of batched KPM against 4.1 s
        with row chunks, tile build included) while the stack's vectors stay
        below the 4 GiB its staging copies address, row chunks (16-entry chunks
        broadcast in 16-lane rows) for other 16- and 32-vector blocks, else row
        groups."""
        if R == 32 and self.precision == L.PREC_FP64 and rows * R * 8 < (1 << 32):
            return L.BATCH_SELL
        return L.BATCH_ROW_CHUNKS if R in (16, 32) else L.BATCH_ROW_GROUPS

    def _geom(self, i):
        g = self._geoms.get(i)
        if g is None:
            g = L.Geometry(L.Deck(W.tblg_cell_deck(i + 1, i, M=self.M)))
            g._origin = origin_orbitals(g)
            th = g.info.angle[1]
            # sheet-2 origin cell moves with the shift: its stacking is (I + X) s
            g._X = np.linalg.solve(LAT.T, (np.eye(2) - rot(-th)) @ rot(th) @ LAT.T)
            self._geoms[i] = g
        return g

    def _run(self, jobs, R):
        """jobs: list of (cell i, shift, strain, targets[R]) -> moments [len(jobs)][R][M]."""
        out = np.zeros((len(jobs), R, self.M))
        order = sorted(range(len(jobs)), key=lambda q: jobs[q][0])   # one geometry upload per cell
        dev = self.dev
        dev.batch_clear()
        pending, targets = [], []
        done = [0]
        rows = [0]                 # rows of the pending stack
        cur = None

        def flush():
            if not pending:
                return
            mode = self.mode if self.mode is not None else self.batch_mode(R, rows[0])
            mu = dev.batch_kpm(self.M, R, targets=np.array(targets), precision=self.precision, graph=False, mode=mode)
            out[np.array(pending)] = mu
            done[0] += len(pending)
            if self.progress:
                self.progress(done[0], len(jobs))
            dev.batch_clear()
            pending.clear()
            targets.clear()
            rows[0] = 0

        for q in order:
            i, sh, e, tg = jobs[q]
            g = self._geom(i)
            if i != cur:
                dev.set_geometry(g)
                dev.pairs_build()
                cur = i
            job = g.jobs()[0]
            job.shift[1][0], job.shift[1][1], job.shift[1][2] = float(sh[0]), float(sh[1]), 0.0
            job.uniform_strain = 1
            job.u[0][0], job.u[0][1], job.u[1][0], job.u[1][1] = e[0], e[2], e[2], e[1]
            job.complex_matrix = 0
            job.energy_rescale, job.energy_shift = RESCALE, 0.0
            dev.assemble(job, L.PATTERN_PRUNED)
            dev.batch_append()
            rows[0] += g.n
            pending.append(q)
            targets.append(np.asarray(tg, dtype=np.int32))
            if len(pending) >= self.chunk:
                flush()
        flush()
        return out

    def samples(self, n_jobs, rng, R=32, theta_range=(0.8, 3.0), strain_sd=0.003, log_theta=True):
        """R random orbitals of each of n_jobs random cells (twist, shift of
        sheet 2, strain).  Returns dict(cfg [n][6], block [n], moments [n][M],
        cell, shift, target) with n = n_jobs R."""
        cfg0 = sample_configs(n_jobs, rng, theta_range=theta_range, strain_sd=strain_sd, log_theta=log_theta)
        jobs, cfg, block, cell, shift, target = [], [], [], [], [], []
        for q in range(n_jobs):
            i = nearest_cell(cfg0[q, 5])
            g = self._geom(i)
            tg = rng.choice(g.n, size=R, replace=False).astype(np.int32)
            b, blk = local_shifts(g, cfg0[q, :2])
            th = abs(math.degrees(g.info.angle[1]))
            for t in tg:
                cfg.append([b[t, 0], b[t, 1], cfg0[q, 2], cfg0[q, 3], cfg0[q, 4], th])
                block.append(blk[t])
                cell.append(i)
                shift.append(cfg0[q, :2])
                target.append(t)
            jobs.append((i, cfg0[q, :2], cfg0[q, 2:5], tg))
        mom = self._run(jobs, R).reshape(-1, self.M)
        return {"cfg": np.array(cfg), "block": np.array(block, dtype=np.int32), "moments": mom,
                "cell": np.array(cell), "shift": np.array(shift), "target": np.array(target)}

    def moments_at(self, cfgs, blocks) -> np.ndarray:
        """Exact moments [n][M] of configurations on the given blocks, at an
        origin-cell orbital of the cell nearest each twist, sheet 2 shifted so
        that the orbital's stacking is b."""
        cfgs = np.atleast_2d(np.asarray(cfgs, dtype=np.float64))
        blocks = np.asarray(blocks)
        jobs = []
        for c, k in zip(cfgs, blocks):
            i = nearest_cell(c[5])
            g = self._geom(i)
            sh = c[:2] if k < 2 else np.linalg.solve(np.eye(2) + g._X, c[:2])
            jobs.append((i, sh, c[2:5], g._origin))
        mu = self._run(jobs, 4)
        return mu[np.arange(len(cfgs)), blocks]


def origin_orbitals(geom) -> np.ndarray:
    """Orbitals of the unit cells at the origin of a periodic bilayer geometry,
    ordered sheet 1 A, B, sheet 2 A, B: the orbitals whose (unshifted)
    positions equal tau_l (sheet 1) and R(theta_2) tau_l (sheet 2) modulo the
    supercell, found by minimum-image distance."""
    i2g, i2p = geom.arrays()
    grid, pos = i2g.reshape(-1, 4), i2p.reshape(-1, 3)[:, :2]
    info = geom.info
    sc = np.array(info.supercell)                 # rows = supercell vectors
    th2 = info.angle[1]
    R2 = np.array([[math.cos(th2), -math.sin(th2)], [math.sin(th2), math.cos(th2)]])
    tau = [np.zeros(2), np.array([0.5 * GR_A, GR_A / (2.0 * math.sqrt(3.0))])]
    want = [(0, 0, tau[0]), (0, 1, tau[1]), (1, 0, R2 @ tau[0]), (1, 1, R2 @ tau[1])]
    out = []
    for s, l, p in want:
        cand = np.nonzero((grid[:, 3] == s) & (grid[:, 2] == l))[0]
        d = pos[cand] - p
        f = np.linalg.solve(sc.T, d.T).T           # supercell coordinates
        f -= np.round(f)
        dist = np.linalg.norm(f @ sc, axis=1)
        k = int(np.argmin(dist))
        if dist[k] > 1e-6:
            raise ValueError(f"no orbital of sheet {s + 1}, sublattice {l} at the origin cell")
        out.append(int(cand[k]))
    return np.array(out, dtype=np.int32)


def train_ensemble(x, t, members=4, hidden=256, layers=3, epochs=60, batch=256, lr=1e-3, seed=0, device=0):
    """Train `members` networks from different initialisations and shuffles.

    The moments vary little between configurations next to their mean, so the
    networks learn the standardised targets (t - mean) / std per output; the
    output layer is then rescaled (W <- std W, b <- std b + mean) so that the
    trained networks, and their ensemble spread, speak in moments again."""
    t = np.asarray(t, dtype=np.float32)
    mu = t.mean(axis=0)
    sd = np.maximum(t.std(axis=0), 1e-6).astype(np.float32)
    ts = ((t - mu) / sd).astype(np.float32)
    nets = [L.Mlp(device, x.shape[1], hidden, layers, t.shape[1], seed=seed * 1000 + e) for e in range(members)]
    # trained side by side; each member's updates are those of Mlp.train with its seed
    losses = list(L.mlp_train_ensemble(nets, x, ts, [seed * 1000 + 100 + e for e in range(members)], epochs=epochs,
                                       batch=batch, lr=lr))
    for net in nets:
        p = net.params()
        dims = net.dims
        nw = dims[-1] * dims[-2]
        w0 = p.size - nw - dims[-1]
        W = p[w0:w0 + nw].reshape(dims[-1], dims[-2])
        b = p[w0 + nw:]
        p[w0:w0 + nw] = (W * sd[:, None]).ravel()
        p[w0 + nw:] = b * sd + mu
        net.set_params(p)
    return nets, losses


def predict_with_fallback(nets, exact: ExactLocalKPM, cfgs, blocks, threshold=1e-3):
    """Ensemble prediction [n][M] of each site's block, with exact local KPM
    where the ensemble spread exceeds the threshold.  Returns (moments,
    fallback mask, spread, report)."""
    t0 = time.perf_counter()
    x = site_features(cfgs, blocks)
    mean, spread = L.mlp_ensemble(nets, x)
    t_inf = time.perf_counter() - t0
    mom = mean.astype(np.float64)
    fb = spread > threshold
    report = {"sites": int(len(cfgs)), "fallback": int(fb.sum()), "inference_s": t_inf}
    if fb.any():
        t1 = time.perf_counter()
        ex = exact.moments_at(cfgs[fb], np.asarray(blocks)[fb])
        report["fallback_s"] = time.perf_counter() - t1
        err = np.abs(mom[fb] - ex)
        report["mlp_max_abs_err_at_fallback"] = float(err.max())
        report["mlp_rms_err_at_fallback"] = float(np.sqrt(np.mean(err ** 2)))
        mom[fb] = ex
    return mom, fb, spread, report


def flake_site_configs(geom, theta_deg: float):
    """Configuration and block of every orbital of a twisted flake (sheet 1 at
    angle 0, sheet 2 at theta): the stacking at its unit cell, no strain."""
    b, block = local_shifts(geom)
    cfg = np.column_stack([b, np.zeros((len(b), 3)), np.full(len(b), theta_deg)])
    return cfg, block
