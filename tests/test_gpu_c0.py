"""GPU parity on the small case as configured: the (7, 6) cell's twist (5.086
deg), about 8,128 orbitals, GrapheneSimple couplings (pruned pattern), real H,
Gershgorin interval with 1 % padding, M = 500 moments of R = 16 random-phase
vectors (seed 0).  The reference expresses it in two ways, and both run here:

  * "flake": an open disk bilayer flake at that twist (MAX_R = 58.54 A,
    8,108 orbitals), the reference's form of an open geometry;
  * "twin": the (7, 6) cell tiled 4 x 4 with periodic boundaries
    (SUPERCELL_M_N = 28 24, 8,128 orbitals).

Bars:
  * Gershgorin interval within 1e-12 (relative) of the one computed from the
    oracle's CSR; rescaled CSR pattern bit-exact, values within 1e-12 relative;
  * FP64 moments within 1e-12 mu_0 of the oracle's faithful single-vector KPM,
    in the doubling scheme and in the reference's scheme (start vectors kept,
    mu_j = Re <v|T_j v>, computeDosTraceKPM);
  * mixed precision (FP32 recursion, FP64 sums): moments within 2e-5 mu_0 of
    the FP64 run; DoS per atom within 1e-6 states/eV/atom inside the spectrum
    (|x| <= 0.99 of the rescaled axis) and within 3e-5 in the two outermost
    bins, where the DoS is zero and the reconstruction divides by
    sqrt(1 - x^2) ~ pi/2M, amplifying the FP32 storage rounding of 8,000
    orbitals about 300 times.
"""
import numpy as np
import pytest

import leo2d_amd as L
from leo2d_amd import workloads as W

pytestmark = pytest.mark.gpu

M, R, SEED = W.C0_M, W.C0_R, 0


@pytest.fixture(scope="module")
def dev():
    d = L.Device(0)
    yield d
    d.close()


def _gershgorin(rp, ci, v):
    """(min_i (d_i - r_i), max_i (d_i + r_i)): d_i the real parts of row i's
    diagonal slots added, r_i its other slots added by modulus."""
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    diag = rows == ci
    d = np.bincount(rows[diag], weights=np.real(v[diag]), minlength=n)
    r = np.bincount(rows[~diag], weights=np.abs(v[~diag]), minlength=n)
    return float(np.min(d - r)), float(np.max(d + r))


@pytest.mark.parametrize("form", ["flake", "twin"])
def test_c0_as_configured(dev, oracle_mod, form):
    print("case c0", form, flush=True)
    text = W.c0_flake_deck() if form == "flake" else W.c0_twin_deck()
    geom = L.Geometry(L.Deck(text))
    assert geom.n == (8108 if form == "flake" else 8128)
    dev.set_geometry(geom)
    dev.pairs_build()
    job, a, b, nnz = W.assemble_rescaled(dev, geom, None)
    rp, ci, v = dev.csr_copy()
    assert not np.iscomplexobj(v)

    od = oracle_mod.parse_deck(text)
    oracle_mod.setup_supercell(od)
    og = oracle_mod.Geometry(oracle_mod.geom_desc(od))
    ojobs = oracle_mod.build_jobs(od, oracle_mod.target_list(od, og))
    oj = oracle_mod.make_job(od, ojobs[0], pruned=1)
    assert oj.complex_matrix == 0
    oj.energy_rescale, oj.energy_shift = 1.0, 0.0
    orp, oci, ov = og.assemble(oj)
    lo, hi = _gershgorin(orp, oci, ov)
    assert abs(a - 1.01 * (hi - lo) / 2.0) <= 1e-12 * a
    assert abs(b - (hi + lo) / 2.0) <= 1e-12 * a
    oj.energy_rescale, oj.energy_shift = a, -b
    orp, oci, ov = og.assemble(oj)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    assert np.all(np.abs(v - ov) <= 1e-12 * np.abs(ov) + 1e-14 * float(np.max(np.abs(ov))))

    n = geom.n
    vecs = oracle_mod.random_phase(n, R, SEED)
    mu64 = dev.kpm(M, random_phase=(R, SEED), scheme=L.SCHEME_DOUBLING)
    print("kernel", dev.timing().kernel, flush=True)
    omu = oracle_mod.kpm(orp, oci, ov, M, vecs=vecs, doubling_all=True, nthreads=16)
    dmu = float(np.max(np.abs(mu64 - omu)))
    print("FP64 max|dmu|/mu0 = %.3e" % (dmu / n), flush=True)
    assert np.max(np.abs(mu64[:, 0] - n)) <= 1e-9
    assert dmu <= 1e-12 * n
    mur = dev.kpm(M, random_phase=(R, SEED), scheme=L.SCHEME_REFERENCE)
    omur = oracle_mod.kpm(orp, oci, ov, M, vecs=vecs, nthreads=16)
    assert np.max(np.abs(mur - omur)) <= 1e-12 * n

    mu32 = dev.kpm(M, random_phase=(R, SEED), scheme=L.SCHEME_DOUBLING, precision=L.PREC_MIXED)
    assert np.max(np.abs(mu32 - mu64)) <= 2e-5 * n
    E, r64 = dev.dos(mu64.mean(axis=0) / n, a, b)
    _, r32 = dev.dos(mu32.mean(axis=0) / n, a, b)
    dd = np.abs(r32[0] - r64[0])
    inner = np.abs((E - b) / a) <= 0.99
    print("mixed max|dDoS| = %.3e states/eV/atom (all bins), %.3e (|x| <= 0.99)" % (dd.max(), dd[inner].max()),
          flush=True)
    assert dd[inner].max() <= 1e-6
    assert dd.max() <= 3e-5
