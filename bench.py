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
    ap.add_argument("--block", type=int, default=0, help="vectors per device block (0 = all R at once)")
    ap.add_argument("--workload", choices=["c0", "c1", "c2", "c3", "c4", "cond"], default="c1",
                    help="c0: the small 5.09 deg case (open flake and periodic 4x4 twin) with the whole CPU "
                         "reference run beside it; c1: 1.05 deg 8x8 tBLG (headline); c2: 32 uneven cells x 8 k-points batched; "
                         "c3: large heterostrained flake, FP32 recursion; c4: learned local-moment surrogate; cond: KPM conductivity moments on a flake")
    ap.add_argument("--cond-radius", type=float, default=200.0, help="cond: flake radius (A)")
    ap.add_argument("--cond-M", type=int, default=512, help="cond: moments per direction")
    ap.add_argument("--c3-radius", type=float, default=1409.0, help="c3: flake radius (A); 1409 A -> 4.7 M orbitals")
    ap.add_argument("--c4-train", type=int, default=200000, help="c4: training samples (orbitals)")
    ap.add_argument("--c4-M", type=int, default=1024, help="c4: local moments per sublattice")
    ap.add_argument("--c4-epochs", type=int, default=100)
    ap.add_argument("--c4-theta-min", type=float, default=0.3,
                    help="c4: smallest training twist (deg); the inference flake's twist must lie in the range")
    ap.add_argument("--c4-radius", type=float, default=650.0, help="c4: inference flake radius (A)")
    ap.add_argument("--c4-theta", type=float, default=0.3, help="c4: inference flake twist (deg)")
    ap.add_argument("--c4-fallback-sample", type=int, default=32,
                    help="c4: flagged sites computed exactly to score the networks and time the fallback")
    ap.add_argument("--mixed-error-Ms", type=str, default="",
                    help="c1 only: instead of the bench line, print max|dDoS| of the FP32 recursion against FP64 "
                         "for each of these comma-separated moment counts (e.g. 1024,2048,4096,8192)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 2 GiB device copy (HBM counter calibration under rocprofv3 --pmc) and "
                         "time it, with a 2 GiB read-only sum and a 2 GiB fill: what the HBM delivers to plain streams")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        x = torch.ones(1 << 28, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(2):
            y.copy_(x)
        torch.cuda.synchronize()
        nb = x.numel() * 8

        def rate(fn, moved, reps=10):
            fn()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(reps):
                fn()
            b.record()
            torch.cuda.synchronize()
            return moved * reps / (a.elapsed_time(b) * 1e-3) / 1e9

        print(json.dumps({"copy_bytes_read": nb, "copy_bytes_written": nb,
                          "copy_GBs": rate(lambda: y.copy_(x), 2 * nb),
                          "read_sum_GBs": rate(lambda: torch.sum(x), nb),
                          "fill_GBs": rate(lambda: y.fill_(2.0), nb)}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "how the path would shard is a design, not measured here.", file=sys.stderr)
        sys.exit(2)

    import leo2d_amd as L

    if args.workload == "c0":
        run_c0(args)
        return
    if args.workload == "c2":
        run_c2(args)
        return
    if args.workload == "c4":
        run_c4(args)
        return
    if args.workload == "c3":
        run_c3(args)
        return
    if args.workload == "cond":
        run_cond(args)
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

    if args.mixed_error_Ms:
        mixed_error_study(dev, n, nnz, R, a_res, b_cen, [int(x) for x in args.mixed_error_Ms.split(",")], log)
        dev.close()
        return

    def run(precision, steps, warmup, time_steps):
        for _ in range(warmup):
            dev.kpm(M, random_phase=(R, 0), scheme=L.SCHEME_DOUBLING, precision=precision, grid=args.grid, block=args.block)
        times = []
        mu = None
        for _ in range(steps):
            t0 = time.perf_counter()
            mu = dev.kpm(M, random_phase=(R, 0), scheme=L.SCHEME_DOUBLING, precision=precision, grid=args.grid, block=args.block)
            times.append(time.perf_counter() - t0)
        # per-launch durations (HIP events around every fused step, separate pass)
        dev.kpm(M, random_phase=(R, 0), scheme=L.SCHEME_DOUBLING, precision=precision, grid=args.grid, block=args.block,
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


def mixed_error_study(dev, n, nnz, R, a_res, b_cen, Ms, log):
    """DoS difference of the FP32 recursion against FP64 on the same job, random
    phases and seed, for several moment counts M: the largest over all M
    energies, its position x = (E - b)/a on the rescaled axis, the largest over
    |x| <= 0.99 and over |x| <= 0.9, and the moment difference relative to mu_0."""
    import leo2d_amd as L

    rows = []
    for M in Ms:
        t0 = time.perf_counter()
        mu64 = dev.kpm(M, random_phase=(R, 0), scheme=L.SCHEME_DOUBLING)
        t64 = time.perf_counter() - t0
        t0 = time.perf_counter()
        mu32 = dev.kpm(M, random_phase=(R, 0), scheme=L.SCHEME_DOUBLING, precision=L.PREC_MIXED)
        t32 = time.perf_counter() - t0
        E, r64 = dev.dos(mu64.mean(axis=0) / n, a_res, b_cen)
        _, r32 = dev.dos(mu32.mean(axis=0) / n, a_res, b_cen)
        r64, r32 = r64[0], r32[0]
        x = (E - b_cen) / a_res
        d = np.abs(r32 - r64)
        k = int(np.argmax(d))
        row = {"M": M, "max_dDoS_all": float(d.max()), "x_at_max": float(x[k]),
               "max_dDoS_abs_x_le_0.99": float(d[np.abs(x) <= 0.99].max()),
               "max_dDoS_abs_x_le_0.9": float(d[np.abs(x) <= 0.9].max()),
               "max_dmu_over_mu0": float(np.max(np.abs(mu32 - mu64)) / n),
               "fp64_s": t64, "mixed_s": t32, "speedup": t64 / t32}
        log(json.dumps(row))
        rows.append(row)
    print(json.dumps({"mixed_error_vs_M": rows, "orbitals": n, "nnz": nnz, "R": R,
                      "units": "states/eV/atom"}), flush=True)


def run_c0(args):
    """The small case: the (7, 6) cell's twist (5.086 deg), about 8,128
    orbitals, GrapheneSimple couplings (pruned), real H, Gershgorin rescale
    with 1 % padding, M = 500, R = 16 random-phase vectors (seed 0), in both
    reference-expressible forms: an open disk flake (8,108 orbitals, the JSON
    line's value) and the (7, 6) cell tiled 4 x 4 periodically (8,128).  One
    "step" is one whole KPM run (250 fused steps and their reductions, the
    moments copied back).  The CPU leg runs the WHOLE workload through the
    reference's own loop around its own SpMatrix::vectorMultiply (oracle/_ref):
    once as the reference runs it under mpirun -n 2 (the root farms jobs, one
    worker computes every vector in turn, locality.cpp:2694-2771, 3103-3115),
    once over every usable core.  The reference keeps real start vectors for a
    real H (computeDosTraceKPM, locality.cpp:6151-6160), so each complex
    random-phase vector v = a + i b runs there as the two real vectors a and b:
    Re <v|T_j(H)|v> = <a|T_j|a> + <b|T_j|b> for a real symmetric H."""
    import ctypes as C

    import leo2d_amd as L
    from leo2d_amd import workloads as W
    from oracle import oracle as orc

    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    M, R, seed = W.C0_M, W.C0_R, 0
    dev = L.Device(0)
    forms = {}
    for form, text in (("flake", W.c0_flake_deck(M)), ("twin", W.c0_twin_deck(M))):
        geom = L.Geometry(L.Deck(text))
        dev.set_geometry(geom)
        t0 = time.perf_counter()
        dev.pairs_build()
        t_pairs = time.perf_counter() - t0
        t0 = time.perf_counter()
        job, a_res, b_cen, nnz = W.assemble_rescaled(dev, geom, None)
        t_fill = time.perf_counter() - t0
        n = geom.n
        res = {"orbitals": n, "nnz": nnz, "rescale_a": a_res, "shift_b": b_cen,
               "pair_search_ms": t_pairs * 1e3, "gershgorin_and_fill_ms": t_fill * 1e3}
        mus = {}
        for prec, name in ((L.PREC_FP64, "fp64"), (L.PREC_MIXED, "mixed")):
            for _ in range(max(1, args.warmup)):
                dev.kpm(M, random_phase=(R, seed), scheme=L.SCHEME_DOUBLING, precision=prec)
            ts = []
            for _ in range(args.steps):
                t0 = time.perf_counter()
                mus[name] = dev.kpm(M, random_phase=(R, seed), scheme=L.SCHEME_DOUBLING, precision=prec)
                ts.append(time.perf_counter() - t0)
            dev.kpm(M, random_phase=(R, seed), scheme=L.SCHEME_DOUBLING, precision=prec, time_steps=True)
            tm = dev.timing()
            t = float(np.mean(ts))
            vb = 16 if prec == L.PREC_FP64 else 8
            sb = step_bytes(n, nnz, R, False, vb, 8 if prec == L.PREC_FP64 else 4)
            res[name] = {"ms_per_run": t * 1e3, "value": nnz * R * (M // 2) / t, "launch_us_avg": tm.step_ms_avg * 1e3,
                         "kernel": kernel_name(tm.kernel, "double2" if vb == 16 else "float2"),
                         "bytes_per_launch": sb, "achieved_GBs": sb / (tm.step_ms_avg * 1e-3) / 1e9}
        E, r64 = dev.dos(mus["fp64"].mean(axis=0) / n, a_res, b_cen)
        _, r32 = dev.dos(mus["mixed"].mean(axis=0) / n, a_res, b_cen)
        dd = np.abs(r32[0] - r64[0])
        inner = np.abs((E - b_cen) / a_res) <= 0.99
        res["mixed"]["speedup_vs_fp64"] = res["fp64"]["ms_per_run"] / res["mixed"]["ms_per_run"]
        res["mixed"]["max_abs_dDoS_states_per_eV_atom"] = float(dd.max())
        res["mixed"]["max_abs_dDoS_inside_spectrum"] = float(dd[inner].max())
        log(f"c0 {form}: n={n} nnz={nnz} a={a_res:.4f}; FP64 {res['fp64']['ms_per_run']:.2f} ms/run "
            f"({res['fp64']['value']:.3e} nnz.vec/s, step launch {res['fp64']['launch_us_avg']:.1f} us), mixed "
            f"{res['mixed']['ms_per_run']:.2f} ms/run, max|dDoS| {dd.max():.2e} (|x|<=0.99: {dd[inner].max():.2e})")

        # the whole workload on the CPU
        rp, ci, v = dev.csr_copy()
        rp32 = np.ascontiguousarray(rp, dtype=np.int32)
        ci32 = np.ascontiguousarray(ci, dtype=np.int32)
        vv = np.ascontiguousarray(v, dtype=np.float64)
        ph = orc.random_phase(n, R, seed)
        real_vecs = np.ascontiguousarray(np.concatenate([ph.real, ph.imag], axis=0))   # [2R][n]
        cores, aff, quota = usable_cores()
        cpu = {}
        if os.path.exists(orc.REF_PATH):
            lib = C.CDLL(orc.REF_PATH, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
            lib.ref_kpm_time.restype = C.c_double
            lib.ref_kpm_time.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                         C.c_int, C.c_int, C.c_int, C.c_void_p]
            for label, th in (("mpirun -n 2 (one worker)", 1), ("all usable cores", cores)):
                mu2 = np.zeros((2 * R, M))
                dt = lib.ref_kpm_time(n, len(ci32), rp32.ctypes.data, ci32.ctypes.data, vv.ctypes.data, 0,
                                      real_vecs.ctypes.data, 2 * R, M, th, mu2.ctypes.data)
                mu = mu2[:R] + mu2[R:]
                err = float(np.max(np.abs(mu - mus["fp64"])) / n)
                cpu[label] = {"value": nnz * R * (M // 2) / dt, "unit": "nnz·vec updates/s", "cores": th,
                              "kind": "reference", "seconds": dt, "max_rel_dmu_vs_gpu": err}
                log(f"c0 {form} CPU reference, {label}: {th} thread(s), whole run in {dt:.3f}s -> "
                    f"{cpu[label]['value']:.3e} nnz.vec/s; max|dmu|/mu0 vs GPU {err:.1e}")
        else:
            t0 = time.perf_counter()
            mu = orc.kpm(rp, ci, v, M, vecs=ph, doubling_all=True, nthreads=cores)
            dt = time.perf_counter() - t0
            cpu["all usable cores"] = {"value": nnz * R * (M // 2) / dt, "unit": "nnz·vec updates/s", "cores": cores,
                                       "kind": "port", "seconds": dt,
                                       "max_rel_dmu_vs_gpu": float(np.max(np.abs(mu - mus["fp64"])) / n)}
        res["cpu"] = cpu
        res["host"] = {"nproc": os.cpu_count(), "affinity": aff, "cgroup_quota": quota, "usable_cores": cores}
        forms[form] = res
    dev.close()

    f = forms["flake"]
    base = f["cpu"].get("mpirun -n 2 (one worker)", f["cpu"].get("all usable cores"))
    out = {
        "metric": METRIC, "value": f["fp64"]["value"], "unit": "nnz·vec updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": f["fp64"]["ms_per_run"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "complex128",
        "data": "synthetic: tBLG geometry built in-process, random-phase start vectors (seed 0)",
        "config": {"workload": "tBLG 5.086 deg ((7,6) cell twist), open disk flake r = 58.54 A (8,108 orbitals), "
                               "GrapheneSimple, real H, Gershgorin rescale (1 % pad), M=500, R=16 random-phase "
                               "vectors, FP64; the periodic (7,6) x 4 x 4 twin beside it",
                   "orbitals": f["orbitals"], "nnz": f["nnz"], "pattern": "pruned", "M": M, "R": R},
        "roofline": {"bound": "hbm", "achieved": f["fp64"]["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": f["fp64"]["achieved_GBs"] / HBM_PEAK_GBS, "traffic": None, "kernel": f["fp64"]["kernel"],
                     "note": "11 MB per step stays in the caches; launch latency, not HBM, bounds this size"},
        "cpu_baseline": dict(base, sample="the whole workload (16 vectors x 250 steps, as 32 real vectors) through "
                                          "the reference's loop around its own SpMatrix::vectorMultiply (oracle/_ref)"),
        "forms": forms,
    }
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
    precs = ((L.PREC_FP64, "fp64"),) if args.no_mixed else ((L.PREC_FP64, "fp64"), (L.PREC_MIXED, "mixed"))
    for prec, name in precs:
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
    mu64 = out_prec["fp64"].pop("mu")
    dmu = None
    if "mixed" in out_prec:
        mu32 = out_prec["mixed"].pop("mu")
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
        "batched": f64, "mixed": out_prec.get("mixed"), "max_rel_dmu_mixed": dmu,
        "one_job_at_a_time": {"value": seq, "kpm_ms": done * 1e3, "speedup_batched": f64["value"] / seq},
        "assembly_s": t_asm,
    }), flush=True)


def run_c3(args):
    """Large strained twisted bilayer flake: theta = 0.5 deg disk of the given
    radius, 0.2 % uniaxial heterostrain on the top sheet with Poisson ratio
    0.16, nearest-neighbour hopping t exp(-3.37 (d/a - 1)) at the strained bond
    length, M = 8192 moments, R = 32 random-phase vectors; FP32 recursion with
    FP64 moment sums, and the FP64 recursion for the DoS error."""
    import leo2d_amd as L
    from leo2d_amd import workloads as W

    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    M, R = 8192, 32
    t0 = time.perf_counter()
    geom = L.Geometry(L.Deck(W.tblg_flake_deck(0.5, args.c3_radius, M=M, material="GrapheneSimpleStrain",
                                               heterostrain=(0.002, 0.16), rescale=1.0)))
    n = geom.n
    dev = L.Device(0)
    dev.set_geometry(geom)
    ni, nx, ns = dev.pairs_build()
    job = geom.jobs()[0]
    dev.assemble(job, L.PATTERN_PRUNED)
    lo, hi = dev.spectral_bounds()
    a_res, b_cen = 1.01 * (hi - lo) / 2.0, (hi + lo) / 2.0
    job.energy_rescale, job.energy_shift = a_res, -b_cen
    nnz = dev.assemble(job, L.PATTERN_PRUNED)
    t_setup = time.perf_counter() - t0
    log(f"c3: {n} orbitals, nnz {nnz} ({nnz / n:.2f}/row), pairs {ni}+{nx}, Gershgorin [{lo:.3f}, {hi:.3f}] eV, "
        f"setup {t_setup:.1f}s")
    res = {}
    for prec, name in ((L.PREC_FP64, "fp64"), (L.PREC_MIXED, "mixed")):
        t1 = time.perf_counter()
        mu = dev.kpm(M, random_phase=(R, 0), scheme=L.SCHEME_DOUBLING, precision=prec)
        t = time.perf_counter() - t1
        _, rho = dev.dos(mu.mean(axis=0) / n, a_res, b_cen)
        res[name] = (t, rho, dev.timing())
        log(f"c3 {name}: {t:.2f}s, {nnz * R * (M // 2) / t:.3e} nnz.vec/s, kernel {dev.timing().kernel}")
    ddos = float(np.max(np.abs(res["mixed"][1] - res["fp64"][1])))
    t64, t32 = res["fp64"][0], res["mixed"][0]
    log(f"c3: mixed {t64 / t32:.2f}x FP64, max|dDoS| = {ddos:.2e} states/eV/atom")
    dev.close()
    upd = nnz * R * (M // 2)
    print(json.dumps({
        "metric": METRIC, "value": upd / t32, "unit": "nnz·vec updates/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": t32 * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "complex64 recursion, f64 moment sums",
        "data": "synthetic: strained flake built in-process, random-phase start vectors (seed 0)",
        "config": {"workload": f"0.5 deg tBLG disk, radius {args.c3_radius} A, 0.2% uniaxial heterostrain (nu 0.16) "
                               "on the top sheet, t exp(-3.37(d/a-1)) nearest-neighbour hopping, M=8192, R=32",
                   "orbitals": n, "nnz": nnz, "pattern": "pruned"},
        "fp64": {"value": upd / t64, "seconds": t64}, "mixed_speedup": t64 / t32,
        "max_abs_dDoS_states_per_eV_atom": ddos, "setup_s": t_setup,
    }), flush=True)


def run_cond(args):
    """KPM conductivity (OBSERVABLE_TYPE = COND): 1.05 deg bilayer flake of the
    given radius (GrapheneSimple, real H), C_nm = (dxH T_n alpha_0).(T_m e_t)
    for the four centre orbitals of both sheets, M x M per target, then the 2D
    Jackson/DCT-III transform.  CPU baseline: the oracle's faithful
    computeCondKPM (M^2 sequential dot products per target) for one target."""
    import leo2d_amd as L
    from leo2d_amd import workloads as W

    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    M = args.cond_M
    geom = L.Geometry(L.Deck(W.tblg_flake_deck(1.05, args.cond_radius, M=M, rescale=10.5)))
    n = geom.n
    dev = L.Device(0)
    dev.set_geometry(geom)
    dev.pairs_build()
    job = geom.jobs()[0]
    nnz = dev.assemble_velocity(job, L.PATTERN_PRUNED)
    targets = [int(geom.info.sheet_count[0]) // 2, int(geom.info.sheet_count[0]) // 2 + 1,
               int(geom.info.sheet_count[0]) + int(geom.info.sheet_count[1]) // 2,
               int(geom.info.sheet_count[0]) + int(geom.info.sheet_count[1]) // 2 + 1]
    R = len(targets)
    dev.cond_moments(M, targets)                     # warm-up
    t0 = time.perf_counter()
    Cm = dev.cond_moments(M, targets)
    t_mom = time.perf_counter() - t0
    tm = dev.cond_timing()
    dev.cond_transform(Cm)                           # warm-up
    t0 = time.perf_counter()
    dev.cond_transform(Cm)
    t_tr = time.perf_counter() - t0
    gemm_flop = 2.0 * M * M * n * R
    # about 3M sparse steps (M - 1 with H for each vector set, M with dxH for the alpha set), each
    # reading the matrix and three vector blocks, plus the two sets written and re-read once
    rec_bytes = 3 * M * (nnz * 12 + 8 * (n + 1) + 3 * n * R * 8) + 2 * M * n * R * 8
    log(f"cond: {n} orbitals, nnz {nnz}, M={M}, R={R}: moments {t_mom * 1e3:.1f} ms (recursion {tm.recursion_ms:.1f} ms, "
        f"GEMM {tm.gemm_ms:.2f} ms = {gemm_flop / (tm.gemm_ms * 1e-3) / 1e12:.1f} TFLOP/s fp64), transform {t_tr * 1e3:.1f} ms")
    out = {
        "metric": "KPM conductivity moments C_nm per second", "value": R * M * M / t_mom, "unit": "moments/s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": t_mom * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: 1.05 deg bilayer flake built in-process, target starts",
        "config": {"workload": f"1.05 deg tBLG disk, radius {args.cond_radius} A, GrapheneSimple, real H, "
                               f"M={M} per direction, {R} centre targets",
                   "orbitals": n, "nnz": nnz, "pattern": "pruned"},
        "roofline": {"bound": "mfma", "achieved": gemm_flop / (tm.gemm_ms * 1e-3) / 1e12, "peak": 78.6,
                     "unit": "TFLOP/s", "frac": gemm_flop / (tm.gemm_ms * 1e-3) / 1e12 / 78.6, "traffic": None,
                     "kernel": "dgemm_kernel<true,true> (C = Dset^T Bset; HIP events around it and the split sum)"},
        "recursion": {"ms": tm.recursion_ms, "algorithmic_GBs": rec_bytes / (tm.recursion_ms * 1e-3) / 1e9},
        "transform_ms": t_tr * 1e3,
    }
    if not args.no_cpu:
        from oracle import oracle as orc
        rp, ci, v = dev.csr_copy()
        dx = dev.velocity_copy()
        Ms = min(M, 256)
        t0 = time.perf_counter()
        ref = orc.cond(rp, ci, v, dx, Ms, targets[:1])
        dt = time.perf_counter() - t0
        # the reference's work per target scales as M^2 n (dots) + 3 M nnz (SpMVs); scale the sample to M
        work_s = Ms * Ms * n + 3 * Ms * nnz
        work = M * M * n + 3 * M * nnz
        rate = (Ms * Ms) / dt * (work_s / work) * (M * M) / (Ms * Ms)
        err = float(np.max(np.abs(Cm[0, :Ms, :Ms] - ref[0])) / np.max(np.abs(ref[0]))) if Ms == M else None
        out["cpu_baseline"] = {"value": rate, "unit": "moments/s", "cores": 1, "kind": "port",
                               "sample": f"one target at M={Ms} ({dt:.1f} s), rate scaled to M={M} by the "
                                         f"reference's M^2 n + 3 M nnz work", "seconds": dt,
                               "max_rel_dC_vs_gpu": err}
        log(f"cond CPU baseline: 1 target, M={Ms} in {dt:.1f}s -> {rate:.3e} moments/s at M={M}")
    dev.close()
    print(json.dumps(out), flush=True)


def run_c4(args):
    """Learned local-moment surrogate.  Exact local moments of random orbitals
    of random commensurate cells (twist log-uniform in [theta_min, 3] deg,
    random shift of sheet 2, strain N(0, 0.3 %)) from batched KPM with target
    starts; an ensemble of four 72 -> 3 x 256 GELU -> M networks trained on the
    GPU with the BF16 MFMA kernels; then inference on every orbital of a
    twisted flake, with exact local KPM where the ensemble spread exceeds 1e-3."""
    import leo2d_amd as L
    from leo2d_amd import surrogate as S
    from leo2d_amd import workloads as W

    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    M, R = args.c4_M, 32
    n_val = max(1024, args.c4_train // 20)
    n_jobs = (args.c4_train + n_val + R - 1) // R
    rng = np.random.default_rng(4)
    dev = L.Device(0)
    last = [time.perf_counter()]

    def progress(done, total):
        if time.perf_counter() - last[0] > 20 or done == total:
            last[0] = time.perf_counter()
            log(f"c4: exact moments for {done} of {total} cells")

    ex = S.ExactLocalKPM(dev, M, chunk=128, progress=progress)
    t0 = time.perf_counter()
    d = ex.samples(n_jobs, rng, R=R, theta_range=(args.c4_theta_min, 3.0))
    t_gen = time.perf_counter() - t0
    n_all = len(d["block"])
    n_train = n_all - n_val
    log(f"c4: {n_all} exact samples ({n_jobs} cells x {R} orbitals, M={M}) in {t_gen:.1f}s ({n_all / t_gen:.0f}/s)")
    x = S.site_features(d["cfg"], d["block"])
    t = d["moments"].astype(np.float32)
    t0 = time.perf_counter()
    nets, losses = S.train_ensemble(x[:n_train], t[:n_train], members=4, hidden=256, layers=3,
                                    epochs=args.c4_epochs, batch=512, lr=1e-3)
    t_train = time.perf_counter() - t0
    dims = nets[0].dims
    fwd_flops = 2 * sum(dims[l] * dims[l + 1] for l in range(len(dims) - 1))
    train_flops = 3 * fwd_flops * n_train * args.c4_epochs * len(nets)
    log(f"c4: trained 4 x {args.c4_epochs} epochs in {t_train:.1f}s ({train_flops / t_train / 1e12:.1f} TFLOP/s "
        f"bf16 incl. host sequencing), final losses {[round(float(l[-1]), 4) for l in losses]}")
    # held-out error, in moments and in the local DoS (Jackson kernel, a = 10.5 eV, b = 0)
    mean, spread_v = L.mlp_ensemble(nets, x[n_train:])
    exv = d["moments"][n_train:]
    bv = d["block"][n_train:]
    rms = float(np.sqrt(np.mean((mean - exv) ** 2)))
    bm = np.stack([t[:n_train][d["block"][:n_train] == k].mean(axis=0) for k in range(4)])
    base = float(np.sqrt(np.mean((bm[bv] - exv) ** 2)))
    _, rho_p = dev.dos(mean.astype(np.float64), S.RESCALE, 0.0)
    _, rho_e = dev.dos(exv, S.RESCALE, 0.0)
    ddos = np.max(np.abs(rho_p - rho_e), axis=1)
    log(f"c4: held-out rms moment error {rms:.3e} (per-block mean predictor {base:.3e}), "
        f"median max|dLDOS| {np.median(ddos):.3e} states/eV, spread median {np.median(spread_v):.2e}")

    # inference on every orbital of the flake
    t0 = time.perf_counter()
    fgeom = L.Geometry(L.Deck(W.tblg_flake_deck(args.c4_theta, args.c4_radius, M=M)))
    scfg, blk = S.flake_site_configs(fgeom, args.c4_theta)
    t_geo = time.perf_counter() - t0
    t0 = time.perf_counter()
    xs = S.site_features(scfg, blk)
    t_feat = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, sp = L.mlp_ensemble(nets, xs, with_mean=False)
    t_inf = time.perf_counter() - t0
    flagged = np.nonzero(sp > 1e-3)[0]
    log(f"c4: flake {args.c4_theta} deg, {len(scfg)} sites: features {t_feat:.2f}s, ensemble {t_inf:.2f}s "
        f"({len(scfg) / t_inf:.3e} sites/s), spread median {np.median(sp):.2e}, {len(flagged)} above 1e-3")
    fb = {}
    if len(flagged):
        pick = rng.choice(flagged, size=min(args.c4_fallback_sample, len(flagged)), replace=False)
        t0 = time.perf_counter()
        exs = ex.moments_at(scfg[pick], blk[pick])
        t_fb = time.perf_counter() - t0
        ms, _ = L.mlp_ensemble(nets, xs[pick])
        err = np.abs(ms - exs)
        per = t_fb / len(pick)
        fb = {"flagged": int(len(flagged)), "exact_sample": int(len(pick)), "exact_s_per_site": per,
              "projected_fallback_s": per * len(flagged),
              "mlp_max_abs_moment_err_at_sample": float(err.max()),
              "mlp_rms_moment_err_at_sample": float(np.sqrt(np.mean(err ** 2)))}
        log(f"c4: exact local KPM at {args.c4_theta} deg: {per * 1e3:.1f} ms/site on {len(pick)} flagged sites; "
            f"network error there rms {fb['mlp_rms_moment_err_at_sample']:.3e}")
    # the networks' error on random flake sites (flagged or not), and the exact cost per site
    pick = rng.choice(len(scfg), size=min(args.c4_fallback_sample, len(scfg)), replace=False)
    t0 = time.perf_counter()
    exr = ex.moments_at(scfg[pick], blk[pick])
    t_rand = time.perf_counter() - t0
    mr, _ = L.mlp_ensemble(nets, xs[pick])
    err_r = np.abs(mr - exr)
    _, rho_p = dev.dos(mr.astype(np.float64), S.RESCALE, 0.0)
    _, rho_e = dev.dos(exr, S.RESCALE, 0.0)
    random_sites = {"sites": int(len(pick)), "rms_moment_err": float(np.sqrt(np.mean(err_r ** 2))),
                    "max_abs_moment_err": float(err_r.max()),
                    "median_max_dLDOS": float(np.median(np.max(np.abs(rho_p - rho_e), axis=1)))}
    log(f"c4: on {len(pick)} random flake sites the ensemble's rms moment error is "
        f"{random_sites['rms_moment_err']:.3e}, median max|dLDOS| {random_sites['median_max_dLDOS']:.3e} states/eV")
    per_site_exact = fb.get("exact_s_per_site", t_rand / len(pick))
    exact_all = per_site_exact * len(scfg)
    surrogate_total = t_feat + t_inf + fb.get("projected_fallback_s", 0.0)
    for m in nets:
        m.close()
    dev.close()
    print(json.dumps({
        "metric": "learned local-moment surrogate: sites/s and error vs exact local KPM", "value": len(scfg) / t_inf,
        "unit": "sites/s (4-network ensemble inference)", "n_gpus": 1, "higher_is_better": True,
        "dtype": "bf16 weights, fp32 accumulation", "data": "synthetic: configurations and flake built in-process",
        "config": {"workload": f"MLP 72 -> 3x256 GELU -> {M}, ensemble of 4, {n_train} training samples "
                               f"(theta {args.c4_theta_min}-3 deg log-uniform, strain N(0, 0.3%)), inference on a "
                               f"{args.c4_theta} deg flake of radius {args.c4_radius} A", "sites": int(len(scfg))},
        "exact_generation": {"samples": int(n_all), "seconds": t_gen},
        "training": {"seconds": t_train, "epochs": args.c4_epochs, "tflops": train_flops / t_train / 1e12,
                     "final_loss": [float(l[-1]) for l in losses]},
        "heldout": {"rms_moment_err": rms, "per_block_mean_rms": base, "median_max_dLDOS": float(np.median(ddos)),
                    "median_spread": float(np.median(spread_v))},
        "inference": {"features_s": t_feat, "ensemble_s": t_inf, "geometry_s": t_geo},
        "fallback": fb, "random_flake_sites": random_sites,
        "gain": {"exact_local_kpm_all_sites_s": exact_all, "surrogate_with_fallback_s": surrogate_total,
                 "speedup": exact_all / surrogate_total},
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
    if variant == 1:
        return f"kpm_step_wave_kernel<{t},{t},1,8>"
    if variant == 15:
        return f"kpm_step_pair_kernel<{t},8>"
    if variant == 14:
        return f"kpm_step_xchunk_kernel<{t},{t},4>"
    if variant == 20:
        return f"kpm_step_sell_kernel<{t},{t},{1 if t == 'double2' else 2},{4 if t == 'double2' else 2}>"
    if variant == 21:
        return f"kpm_step_sell_cw_kernel<{t},{t},{1 if t == 'double2' else 2},{4 if t == 'double2' else 2}>"
    return f"kpm step variant {variant}"


def usable_cores():
    """Cores this process may run on: the affinity set, capped by the cgroup
    CPU quota when one is set (cpu.max, "quota period")."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, int(float(q) / float(per)))
    except (OSError, ValueError):
        pass
    return (min(aff, quota) if quota else aff), aff, quota


def cpu_baseline(dev, nnz, n, mu_gpu, log):
    """CPU KPM on the same matrix over every usable host core, one start vector
    per core (the reference gives each worker whole vectors, locality.cpp:
    3103-3115):
      * "reference": the reference's own loop around its own
        SpMatrix::vectorMultiply (spmatrix.cpp:346-410, the non-MKL CSR loop),
        compiled from its sources into oracle/_ref -- per step one SpMV,
        three buffers, two copies, one subtract, sequential dots;
      * "port": the oracle's restatement of the same loop (oracle/leo2d_oracle.c).
    A bounded sample (one vector per core, about 12 s) is timed; its moments
    are checked against the GPU's."""
    import ctypes as C

    from oracle import oracle as orc

    rp, ci, v = dev.csr_copy()
    cores, aff, quota = usable_cores()
    log(f"host: nproc {os.cpu_count()}, affinity {aff}, cgroup quota {quota}, usable cores {cores}")
    R_s = cores
    cpx = np.iscomplexobj(v)
    rp32 = np.ascontiguousarray(rp, dtype=np.int32)
    ci32 = np.ascontiguousarray(ci, dtype=np.int32)
    vv = np.ascontiguousarray(v.view(np.float64) if cpx else v)
    have_ref = os.path.exists(orc.REF_PATH)
    # per-SpMV cost on one core sets the sample length (about 12 s of loops)
    steps = 256
    if have_ref:
        L = C.CDLL(orc.REF_PATH, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        L.ref_spmv_time.restype = C.c_double
        L.ref_spmv_time.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                    C.c_void_p, C.c_int]
        L.ref_kpm_time.restype = C.c_double
        L.ref_kpm_time.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                   C.c_int, C.c_int, C.c_int, C.c_void_p]
        x = np.ones(n, dtype=np.complex128 if cpx else np.float64)
        y = np.zeros_like(x)
        t1 = L.ref_spmv_time(n, len(ci), rp32.ctypes.data, ci32.ctypes.data, vv.ctypes.data, int(cpx),
                             x.ctypes.data, y.ctypes.data, 2) / 2
        steps = int(max(16, min(1024, 12.0 / max(t1 * 1.6, 1e-6))))
    steps -= steps % 2          # M = 2 steps, a multiple of 4 (POLY_ORDER % 4, main.cpp:914)
    M_s = 2 * steps
    vecs = orc.random_phase(n, R_s, 0)
    out = {}
    legs = []
    if have_ref:
        mu = np.zeros((R_s, M_s))
        vflat = np.ascontiguousarray(vecs).view(np.float64)
        dt = L.ref_kpm_time(n, len(ci), rp32.ctypes.data, ci32.ctypes.data, vv.ctypes.data, int(cpx),
                            vflat.ctypes.data, R_s, M_s, cores, mu.ctypes.data)
        legs.append(("reference", dt, mu))
    t0 = time.perf_counter()
    mu = orc.kpm(rp, ci, v, M_s, vecs=vecs, doubling_all=True, nthreads=cores)
    legs.append(("port", time.perf_counter() - t0, mu))
    for kind, dt, mu in legs:
        rate = nnz * R_s * steps / dt
        err = float(np.max(np.abs(mu - mu_gpu[:R_s, :M_s])) / n)
        what = ("the reference's loop around its own SpMatrix::vectorMultiply (oracle/_ref)" if kind == "reference"
                else "the oracle's restatement of the loop")
        log(f"CPU {kind}: {cores} cores, {R_s} vectors x {steps} steps in {dt:.1f}s -> {rate:.3e} nnz.vec/s; "
            f"max|dmu|/mu0 vs GPU = {err:.2e}")
        out[kind] = {"value": rate, "unit": "nnz·vec updates/s", "cores": cores, "kind": kind,
                     "sample": f"{R_s} random-phase vectors (one per core) x {steps} Chebyshev steps (M={M_s}) of "
                               f"{what} on the same {n}-orbital complex H",
                     "seconds": dt, "max_rel_dmu_vs_gpu": err}
    best = out.get("reference", out["port"])
    res = dict(best)
    res["host"] = {"nproc": os.cpu_count(), "affinity": aff, "cgroup_quota": quota, "usable_cores": cores}
    if "reference" in out:
        res["port"] = out["port"]
    return res


if __name__ == "__main__":
    main()
