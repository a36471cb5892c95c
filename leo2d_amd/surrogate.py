"""Learned local-moment surrogate with exact local-KPM fallback.

A local stacking configuration c = (b1, b2, exx, eyy, exy, theta) is: the
shift of sheet 2 against sheet 1 in fractional lattice units, a symmetric
strain tensor applied to both sheets, and the twist angle in degrees.  Its
exact local moments are the local-DoS Chebyshev moments of the four orbitals
of the unit cell at the origin (sheet 1 A, B; sheet 2 A, B) in the
commensurate (i+1, i) twisted-bilayer cell whose twist is nearest theta, with
sheet 2 shifted by b (setConfigPositions shifts, src/locality.cpp:3577-3660)
and the strain applied as UNIFORM_STRAIN (src/locality.cpp:3849-3857), on the
GrapheneSimple couplings, rescaled by a fixed a = 10.5 eV.  They come from the
library's batched KPM with target starts (one job per configuration).

An ensemble of MLPs (64 Fourier features -> 3 x 256 GELU -> 4 M moments,
bfloat16 weights, FP32 accumulation, trained on the GPU) predicts them; a
configuration whose ensemble spread (largest standard deviation over the
outputs) exceeds a threshold falls back to the exact computation, which also
scores the networks' error there.
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
    """Exact local moments [n][4][M] of configurations on one device handle."""

    def __init__(self, dev: L.Device, M: int, chunk: int = 512, precision=L.PREC_FP64):
        self.dev, self.M, self.chunk, self.precision = dev, M, chunk, precision
        self._geoms = {}

    def _geom(self, i):
        g = self._geoms.get(i)
        if g is None:
            g = L.Geometry(L.Deck(W.tblg_cell_deck(i + 1, i, M=self.M)))
            g._origin = origin_orbitals(g)
            self._geoms[i] = g
        return g

    def moments(self, cfgs) -> np.ndarray:
        cfgs = np.atleast_2d(np.asarray(cfgs, dtype=np.float64))
        n = cfgs.shape[0]
        out = np.zeros((n, 4, self.M))
        cells = np.array([nearest_cell(t) for t in cfgs[:, 5]])
        order = np.argsort(cells, kind="stable")      # one geometry upload per cell
        dev = self.dev
        dev.batch_clear()
        pending, targets = [], []
        cur = None

        def flush():
            if not pending:
                return
            mu = dev.batch_kpm(self.M, 4, targets=np.array(targets), precision=self.precision, graph=False)
            out[np.array(pending)] = mu
            dev.batch_clear()
            pending.clear()
            targets.clear()

        for q in order:
            i = int(cells[q])
            g = self._geom(i)
            if i != cur:
                dev.set_geometry(g)
                dev.pairs_build()
                cur = i
            b1, b2, exx, eyy, exy, _ = cfgs[q]
            job = g.jobs()[0]
            job.shift[1][0], job.shift[1][1], job.shift[1][2] = b1, b2, 0.0
            job.uniform_strain = 1
            job.u[0][0], job.u[0][1], job.u[1][0], job.u[1][1] = exx, exy, exy, eyy
            job.complex_matrix = 0
            job.energy_rescale, job.energy_shift = RESCALE, 0.0
            dev.assemble(job, L.PATTERN_PRUNED)
            dev.batch_append()
            pending.append(q)
            targets.append(g._origin)
            if len(pending) >= self.chunk:
                flush()
        flush()
        return out


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
    nets, losses = [], []
    for e in range(members):
        net = L.Mlp(device, L.MLP_FEATURES, hidden, layers, t.shape[1], seed=seed * 1000 + e)
        losses.append(net.train(x, ts, epochs=epochs, batch=batch, lr=lr, seed=seed * 1000 + 100 + e))
        p = net.params()
        dims = net.dims
        nw = dims[-1] * dims[-2]
        w0 = p.size - nw - dims[-1]
        W = p[w0:w0 + nw].reshape(dims[-1], dims[-2])
        b = p[w0 + nw:]
        p[w0:w0 + nw] = (W * sd[:, None]).ravel()
        p[w0 + nw:] = b * sd + mu
        net.set_params(p)
        nets.append(net)
    return nets, losses


def predict_with_fallback(nets, exact: ExactLocalKPM, cfgs, threshold=1e-3, blocks=None):
    """Ensemble prediction [n][4][M] with exact local KPM where the spread
    (RMS of the ensemble standard deviation over the M moments of a
    sublattice) exceeds the threshold: over the sublattice blocks[i] of each
    configuration when given, else over its largest of the four.  Returns
    (moments, fallback mask, spread, report)."""
    M = exact.M
    t0 = time.perf_counter()
    x = L.mlp_features(cfgs)
    mean, sp4 = L.mlp_ensemble(nets, x, blocks=4)
    spread = sp4.max(axis=1) if blocks is None else sp4[np.arange(len(x)), blocks]
    t_inf = time.perf_counter() - t0
    mom = mean.reshape(-1, 4, M).astype(np.float64)
    fb = spread > threshold
    report = {"sites": int(len(cfgs)), "fallback": int(fb.sum()), "inference_s": t_inf}
    if fb.any():
        t1 = time.perf_counter()
        ex = exact.moments(cfgs[fb])
        report["fallback_s"] = time.perf_counter() - t1
        err = np.abs(mom[fb] - ex)
        report["mlp_max_abs_err_at_fallback"] = float(err.max())
        report["mlp_rms_err_at_fallback"] = float(np.sqrt(np.mean(err ** 2)))
        mom[fb] = ex
    return mom, fb, spread, report


def flake_site_configs(geom, theta_deg: float):
    """Configuration of every orbital of a twisted flake (sheet 1 at angle 0,
    sheet 2 at theta): the local shift of sheet 2's lattice at the orbital's
    unit cell's A position r, b = A^-1 (I - R(-theta)) r mod 1 in sheet 2's
    lattice basis (the shift that the training cells apply at their origin),
    no strain.  Returns (cfg [n][6], block [n] = 2 sheet + sublattice)."""
    i2g, i2p = geom.arrays()
    grid, pos = i2g.reshape(-1, 4), i2p.reshape(-1, 3)
    th = math.radians(theta_deg)
    Rm = np.array([[math.cos(-th), -math.sin(-th)], [math.sin(-th), math.cos(-th)]])
    R2 = Rm.T
    # the unit cell's A position: r - tau_l on sheet 1, r - R(theta) tau_l on sheet 2
    tauB = np.array([0.5 * GR_A, GR_A / (2.0 * math.sqrt(3.0))])
    off = np.where(grid[:, 2:3] == 1, np.where(grid[:, 3:4] == 1, (R2 @ tauB)[None, :], tauB[None, :]), 0.0)
    r = pos[:, :2] - off
    d = r - r @ Rm.T
    b = np.linalg.solve(LAT.T, d.T).T % 1.0
    cfg = np.column_stack([b, np.zeros((len(r), 3)), np.full(len(r), theta_deg)])
    block = 2 * grid[:, 3] + grid[:, 2]
    return cfg, block.astype(np.int32)
