#!/usr/bin/env python3
"""Benchmark of the KPM hot path on one MI355X.

Workload (the headline configuration): twisted bilayer graphene at 1.05 deg,
commensurate cell (M, N) = (32, 31) tiled 8 x 8 with periodic boundaries, i.e.
SUPERCELL_M_N = 256 248 -> 762,112 orbitals; nearest-neighbour intralayer
t = -2.7 eV, interlayer t_perp(r) = 0.48 exp(-(r - 3.35)/0.45) eV for r < 6 A
(pruned pattern); complex Bloch Hamiltonian H(k) at k = 0.3 b1 + 0.2 b2;
energy rescaling by the Gershgorin interval with 1% padding; M = 2048
Chebyshev moments for R = 64 random-phase start vectors (seed 0).

One "step" = one full KPM run on the assembled Hamiltonian: M/2 = 1024 fused
block-SpMM Chebyshev launches over all 64 vectors plus their moment
reductions.  Metric: nnz x R x (M/2) / t_step (nnz-vector updates per second).
Assembly (GPU pair search + fill) is timed separately.  The CPU baseline is
the oracle's faithful single-vector KPM on the same matrix, timed on a bounded
sample over the host's cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "KPM nnz·vec updates/s & %HBM peak; mixed vs FP64 at ≤1e-6 DoS error"
HBM_PEAK_GBS = 8000.0


def deck_c1(M=2048):
    return f"""JOB_NAME = tblg_1p05_8x8
MAXSIZE = 761
MINSIZE = -761
BOUNDARY_CONDITION = 1
SUPERCELL_M_N = 256 248
NUM_SHEETS = 2
START_SHEET 1
MATERIAL = GrapheneSimple
HEIGHT = 0
ANGLE = 0
END_SHEET 1
START_SHEET 2
MATERIAL = GrapheneSimple
HEIGHT = 3.35
ANGLE = 0
END_SHEET 2
SOLVER_TYPE = SQ
NUM_TARGET_SHEETS = 1
TARGET_SHEETS = 1
NUM_SHIFT_SHEETS = 1
SHIFT_SHEETS = 2
NSHIFTS = 1
ENERGY_RESCALE = 1
ENERGY_SHIFT = 0.0
POLY_ORDER = {M}
"""


def step_bytes(n, nnz, R, cpx_mat, vec_bytes, val_bytes):
    """Algorithmic HBM bytes of one fused step: H once (values + int32 columns
    + int64 row pointers), T_j read once, T_{j-1} read once, T_{j+1} written once."""
    return nnz * (val_bytes + 4) + 8 * (n + 1) + 3 * n * R * vec_bytes


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--M", type=int, default=2048)
    ap.add_argument("--R", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--workload", choices=["c1", "c2"], default="c1",
                    help="c1: 1.05 deg 8x8 tBLG (headline); c2: 32 uneven cells x 8 k-points batched")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 2 GiB device copy (HBM counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        x = torch.ones(1 << 28, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(2):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 8, "copy_bytes_written": y.numel() * 8}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "how the path would shard is a design, not measured here.", file=sys.stderr)
        sys.exit(2)

    import leo2d_amd as L

    if args.workload == "c2":
        run_c2(args)
        return

    M, R = args.M, args.R
    log = lambda *a: print(*a, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    deck = L.Deck(deck_c1(M))
    geom = L.Geometry(deck)
    t_geom = time.perf_counter() - t0
    log(f"geometry: n={geom.n} ({t_geom:.2f}s host)")
    dev = L.Device(0)
    dev.set_geometry(geom)
    t0 = time.perf_counter()
    ni, nx, ns = dev.pairs_build()
    t_pairs = time.perf_counter() - t0
    log(f"pair search: intra={ni} inter={nx} slots={ns} ({t_pairs:.2f}s)")

    job = geom.jobs()[0]
    info = geom.info
    # Bloch vector k = 0.3 b1 + 0.2 b2 of the supercell
    a = np.array(info.supercell)
    A3 = np.array([[a[0, 0], a[0, 1], 0.0], [a[1, 0], a[1, 1], 0.0], [0.0, 0.0, 1.0]])
    B = 2 * np.pi * np.linalg.inv(A3).T
    k = 0.3 * B[0, :2] + 0.2 * B[1, :2]
    job.complex_matrix = 1
    job.k[0], job.k[1] = float(k[0]), float(k[1])
    # Gershgorin interval of the unscaled H, 1% padding
    job.energy_rescale, job.energy_shift = 1.0, 0.0
    t0 = time.perf_counter()
    dev.assemble(job, L.PATTERN_PRUNED)
    t_fill = time.perf_counter() - t0
    lo, hi = dev.spectral_bounds()
    a_res = 1.01 * (hi - lo) / 2.0
    b_cen = (hi + lo) / 2.0
    job.energy_rescale, job.energy_shift = a_res, -b_cen
    t0 = time.perf_counter()
    nnz = dev.assemble(job, L.PATTERN_PRUNED)
    t_fill2 = time.perf_counter() - t0
    n = geom.n
    log(f"assembly: nnz={nnz} ({nnz / n:.2f}/row), Gershgorin [{lo:.4f}, {hi:.4f}] eV, "
        f"a={a_res:.4f} b={b_cen:.4f}, fill {t_fill2 * 1e3:.1f} ms")

    def run(precision, steps, warmup, time_steps):
        for _ in range(warmup):
            dev.kpm(M, random_phase=(R, 0), scheme=L.SCHEME_DOUBLING, precision=precision, grid=args.grid)
        times = []
        mu = None
        for _ in range(steps):
            t0 = time.perf_counter()
            mu = dev.kpm(M, random_phase=(R, 0), scheme=L.SCHEME_DOUBLING, precision=precision, grid=args.grid)
            times.append(time.perf_counter() - t0)
        # per-launch durations (HIP events around every fused step, separate pass)
        dev.kpm(M, random_phase=(R, 0), scheme=L.SCHEME_DOUBLING, precision=precision, grid=args.grid,
                time_steps=time_steps)
        tm = dev.timing()
        return mu, times, tm

    mu64, times64, tm64 = run(L.PREC_FP64, args.steps, args.warmup, True)
    t_step = float(np.mean(times64))
    value = nnz * R * (M // 2) / t_step
    sb = step_bytes(n, nnz, R, True, 16, 16)
    achieved = sb / (tm64.step_ms_avg * 1e-3) / 1e9
    log(f"FP64: {t_step * 1e3:.1f} ms/step, {value:.3e} nnz.vec/s, step launch {tm64.step_ms_avg * 1e3:.1f} us, "
        f"{achieved:.0f} GB/s ({achieved / HBM_PEAK_GBS * 100:.1f}% of 8 TB/s)")

    # DoS per atom (trace estimate): moments averaged over vectors, / n
    E, rho64 = dev.dos(mu64.mean(axis=0) / n, a_res, b_cen)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "nnz·vec updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic: commensurate tBLG geometry built in-process, random-phase start vectors (seed 0)",
        "config": {
            "workload": "tBLG 1.05 deg, (32,31) cell tiled 8x8 periodic (SUPERCELL_M_N 256 248), complex H(k), "
                        "pruned nearest-neighbour + exp interlayer (r<6 A) couplings, M=2048 moments, "
                        "R=64 random-phase vectors, FP64 recursion",
            "orbitals": n, "nnz": nnz, "nnz_per_row": nnz / n, "pattern": "pruned",
            "M": M, "R": R, "gershgorin_eV": [lo, hi], "rescale_a": a_res, "shift_b": b_cen,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel": kernel_name(tm64.kernel, "double2"),
            "bytes_per_launch": sb, "launch_us_avg": tm64.step_ms_avg * 1e3,
        },
        "assembly": {"geometry_host_s": t_geom, "pair_search_s": t_pairs, "fill_ms": t_fill2 * 1e3,
                     "first_fill_ms": t_fill * 1e3},
    }
    out["roofline"]["traffic"] = measured_traffic(out["roofline"]["kernel"], nnz, R)

    if not args.no_mixed:
        mu32, times32, tm32 = run(L.PREC_MIXED, args.steps, 1, True)
        t32 = float(np.mean(times32))
        _, rho32 = dev.dos(mu32.mean(axis=0) / n, a_res, b_cen)
        ddos = float(np.max(np.abs(rho32 - rho64)))
        sb32 = step_bytes(n, nnz, R, True, 8, 8)
        out["mixed"] = {
            "value": nnz * R * (M // 2) / t32, "ms_per_step": t32 * 1e3, "speedup_vs_fp64": t_step / t32,
            "max_abs_dDoS_states_per_eV_atom": ddos, "dtype": "complex64 recursion, f64 moment sums",
            "launch_us_avg": tm32.step_ms_avg * 1e3, "kernel": kernel_name(tm32.kernel, "float2"),
            "traffic": measured_traffic(kernel_name(tm32.kernel, "float2"), nnz, R),
            "achieved_GBs": sb32 / (tm32.step_ms_avg * 1e-3) / 1e9,
        }
        log(f"mixed: {t32 * 1e3:.1f} ms/step ({t_step / t32:.2f}x), max|dDoS| = {ddos:.2e} states/eV/atom")

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(dev, nnz, n, mu64, log)

    dev.close()
    print(json.dumps(out), flush=True)


def run_c2(args):
    """Batched uneven solves: 32 commensurate tBLG cells (i+1, i) with i drawn
    uniformly from 9..31 (seed 1), 8 Bloch k-points per cell drawn uniformly in
    the moire Brillouin zone, 16 random-phase vectors per k, M = 1024, all 256
    jobs stacked and advanced by one hipGraph-captured recursion with the
    autotuned launch shape.  The same jobs run one by one through the
    single-job entry point give the unbatched rate beside it."""
    import leo2d_amd as L
    from leo2d_amd import workloads as W

    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    M, R, cells, kpts = 1024, 16, 32, 8
    rng = np.random.default_rng(1)
    idx = rng.integers(9, 32, size=cells)
    kfr = rng.random((cells, kpts, 2))
    dev = L.Device(0)
    t0 = time.perf_counter()
    nnz_jobs = []
    for c, i in enumerate(idx):
        geom = L.Geometry(L.Deck(W.tblg_cell_deck(int(i) + 1, int(i), M=M)))
        dev.set_geometry(geom)
        dev.pairs_build()
        for k in range(kpts):
            _, _, _, nnz = W.assemble_rescaled(dev, geom, kfr[c, k])
            dev.batch_append()
            nnz_jobs.append(nnz)
    t_asm = time.perf_counter() - t0
    jobs, rows, nnz = dev.batch_info()
    log(f"c2: {jobs} jobs, {rows} rows (cells i = {sorted(int(x) for x in idx)}), nnz {nnz}, assembly {t_asm:.2f}s")
    updates = nnz * R * (M // 2)

    out_prec = {}
    for prec, name in ((L.PREC_FP64, "fp64"), (L.PREC_MIXED, "mixed")):
        dev.batch_kpm(M, R, seed=0, precision=prec, tune=True)        # tune + capture
        tt = dev.batch_timing()
        for _ in range(args.warmup):
            dev.batch_kpm(M, R, seed=0, precision=prec, tune=True)
        times, gpu_ms = [], []
        for _ in range(args.steps):
            t1 = time.perf_counter()
            mu = dev.batch_kpm(M, R, seed=0, precision=prec, tune=True)
            times.append(time.perf_counter() - t1)
            gpu_ms.append(dev.batch_timing().recursion_ms)
        t = float(np.mean(times))
        tb = dev.batch_timing()
        out_prec[name] = {"value": updates / t, "ms_per_step": t * 1e3, "graph_ms": float(np.mean(gpu_ms)),
                          "mode": tb.mode, "tile_rows": tb.tile_rows, "grid": tb.grid,
                          "tune_ms": tt.tune_ms, "candidates": tt.candidates, "mu": mu}
        log(f"c2 {name}: {t * 1e3:.1f} ms per batch ({updates / t:.3e} nnz.vec/s), graph {np.mean(gpu_ms):.1f} ms, "
            f"tuned mode={tb.mode} tile_rows={tb.tile_rows} grid={tb.grid} ({tt.candidates} shapes, {tt.tune_ms:.0f} ms)")

    # the same jobs one at a time through leo2d_kpm_moments (no batching, no graph)
    t0 = time.perf_counter()
    done = 0
    for c, i in enumerate(idx):
        geom = L.Geometry(L.Deck(W.tblg_cell_deck(int(i) + 1, int(i), M=M)))
        dev.set_geometry(geom)
        dev.pairs_build()
        for k in range(kpts):
            W.assemble_rescaled(dev, geom, kfr[c, k])
            t1 = time.perf_counter()
            dev.kpm(M, random_phase=(R, 0), scheme=L.SCHEME_DOUBLING)
            done += time.perf_counter() - t1
    seq = updates / done
    log(f"c2 one job at a time: {done * 1e3:.1f} ms of KPM ({seq:.3e} nnz.vec/s)")
    dev.close()
    mu64, mu32 = out_prec["fp64"].pop("mu"), out_prec["mixed"].pop("mu")
    dmu = float(np.max(np.abs(mu64 - mu32)) / np.max(np.abs(mu64[:, :, 0])))
    f64 = out_prec["fp64"]
    print(json.dumps({
        "metric": METRIC, "value": f64["value"], "unit": "nnz·vec updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": f64["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "complex128",
        "data": "synthetic: 32 commensurate tBLG cells built in-process, random-phase start vectors (seed 0)",
        "config": {"workload": "batched uneven solves: 32 tBLG cells (i+1, i), i ~ U{9..31} (seed 1), 8 k-points "
                               "each, R=16 per k, M=1024, pruned GrapheneSimple H(k), one hipGraph per batch",
                   "jobs": jobs, "rows": rows, "nnz": nnz, "cells_i": [int(x) for x in idx]},
        "batched": f64, "mixed": out_prec["mixed"], "max_rel_dmu_mixed": dmu,
        "one_job_at_a_time": {"value": seq, "kpm_ms": done * 1e3, "speedup_batched": f64["value"] / seq},
        "assembly_s": t_asm,
    }), flush=True)


def measured_traffic(kernel, nnz, R):
    """HBM bytes per launch of `kernel` on this workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by the copy-kernel
    calibration), or None when no pass covers this kernel and shape."""
    path = os.path.join(ROOT, "profiles", "kpm_step_c1_traffic.json")
    try:
        pj = json.load(open(path))
    except (OSError, ValueError):
        return None
    if pj.get("nnz") != nnz or pj.get("R") != R:
        return None
    k = pj.get("kernels", {}).get(kernel)
    return k["bytes_per_launch"] if k else None


def kernel_name(variant, t):
    """Name of the fused step kernel a KPM run launched (complex H, R = 64)."""
    if variant == 8:
        return f"kpm_step_smem_kernel<{t},{t},1,8>"
    if variant == 1:
        return f"kpm_step_wave_kernel<{t},{t},1,8,256>"
    return f"kpm step variant {variant}"


def cpu_baseline(dev, nnz, n, mu_gpu, log):
    """The oracle's faithful single-vector KPM (computeDosKPM-style loop: one
    CSR SpMV per step, 3 buffers, 2 copies, 1 subtract, sequential dots) on the
    same matrix, one vector per worker thread, all workers busy."""
    from oracle import oracle as orc

    rp, ci, v = dev.csr_copy()
    threads = min(16, os.cpu_count() or 1)
    R_s = threads
    M_s = 768   # 384 steps: about 10-15 s of CPU work on 16 threads
    vecs = orc.random_phase(n, R_s, 0)
    t0 = time.perf_counter()
    mu = orc.kpm(rp, ci, v, M_s, vecs=vecs, doubling_all=True, nthreads=threads)
    dt = time.perf_counter() - t0
    steps = M_s // 2
    rate = nnz * R_s * steps / dt
    err = float(np.max(np.abs(mu - mu_gpu[:R_s, :M_s])) / n)
    log(f"CPU baseline: {threads} threads, {R_s} vectors x {steps} steps in {dt:.1f}s -> {rate:.3e} nnz.vec/s; "
        f"max|dmu|/mu0 vs GPU = {err:.2e}")
    return {"value": rate, "unit": "nnz·vec updates/s", "cores": threads, "kind": "port",
            "sample": f"{R_s} random-phase vectors x {steps} Chebyshev steps (M={M_s}) on the same "
                      f"{n}-orbital complex H, one vector per thread",
            "seconds": dt, "max_rel_dmu_vs_gpu": err}


if __name__ == "__main__":
    main()
