"""GPU parity on the headline workload's own cell.

The 1.05 deg (32, 31) twisted-bilayer cell (11,908 orbitals), untiled and tiled
2 x 2 (47,632 orbitals), periodic, GrapheneSimple couplings (pruned pattern),
complex H(k) at k = 0.3 b1 + 0.2 b2, Gershgorin interval with 1 % padding -- the
construction bench.py uses for the 8 x 8 tiling -- then M = 2048 moments of
R = 64 random-phase vectors (seed 0) with the production step kernel.

Bars:
  * Gershgorin interval of the unscaled H(k) within 1e-12 (relative) of the
    one computed on the host from the oracle's CSR;
  * rescaled CSR: pattern bit-exact, values |dv| <= 1e-12 |v| + 1e-14 max|H|;
  * FP64 moments within 1e-12 mu_0 of the oracle's faithful single-vector KPM
    (computeDosKPM loop) on the first 16 of the 64 vectors;
  * mixed precision (FP32 recursion, FP64 moment sums) against the FP64 run:
    moments within 2e-5 mu_0; DoS per atom (trace over the 64 vectors, Jackson
    kernel) within 1e-6 states/eV/atom at every energy on the full 8 x 8 cell
    (762,112 orbitals, the size the bar is stated for), and inside the spectrum
    (|x| <= 0.99 of the rescaled axis) on the small cells.

On the small cells the largest DoS difference sits in the two outermost
energy bins, x = -cos(pi/2M) and +cos(pi/2M), where the DoS itself is zero
(outside the Gershgorin-padded spectrum) and the reconstruction divides by
sqrt(1 - x^2) ~ pi/2M: the FP32 rounding of the stored T_j (about 6e-8 per
element per step) summed over every moment is amplified about 1300 times
there.  Its size falls as 1/sqrt(n) with the number of orbitals averaged, so
those bins are bounded at 3e-4 on the 11,908-orbital cell and at 1e-6 on the
full 8 x 8 cell.
"""
import numpy as np
import pytest

import leo2d_amd as L
from leo2d_amd import workloads as W

pytestmark = pytest.mark.gpu

M = 2048
R = 64
K_FRAC = (0.3, 0.2)


@pytest.fixture(scope="module")
def dev():
    d = L.Device(0)
    yield d
    d.close()


def _gershgorin(rp, ci, v):
    """lo = min_i (Re H_ii - sum_{j != i} |H_ij|), hi likewise: the real parts
    of a row's diagonal slots added, its other slots added by modulus."""
    n = len(rp) - 1
    lo, hi = np.inf, -np.inf
    for i in range(n):
        cols = ci[rp[i]:rp[i + 1]]
        vals = v[rp[i]:rp[i + 1]]
        d = np.sum(vals[cols == i]).real
        r = np.sum(np.abs(vals[cols != i]))
        lo = min(lo, d - r)
        hi = max(hi, d + r)
    return lo, hi


@pytest.mark.parametrize("tiles", [1, 2])
def test_c1_cell_assembly_and_moments(dev, oracle_mod, tiles):
    print("case (32,31) x", tiles, flush=True)
    text = W.tblg_cell_deck(32, 31, tiles, M=M)
    geom = L.Geometry(L.Deck(text))
    assert geom.n == 11908 * tiles * tiles
    dev.set_geometry(geom)
    dev.pairs_build()
    job, a, b, nnz = W.assemble_rescaled(dev, geom, K_FRAC)
    rp, ci, v = dev.csr_copy()

    # the oracle's assembly of the same job, unscaled then rescaled
    od = oracle_mod.parse_deck(text)
    oracle_mod.setup_supercell(od)
    og = oracle_mod.Geometry(oracle_mod.geom_desc(od))
    ojobs = oracle_mod.build_jobs(od, oracle_mod.target_list(od, og))
    oj = oracle_mod.make_job(od, ojobs[0], pruned=1)
    oj.complex_matrix = 1
    oj.k[0], oj.k[1] = job.k[0], job.k[1]
    oj.energy_rescale, oj.energy_shift = 1.0, 0.0
    orp, oci, ov = og.assemble(oj)
    lo, hi = _gershgorin(orp, oci, ov)
    pad = 1.01
    assert abs(a - pad * (hi - lo) / 2.0) <= 1e-12 * a
    assert abs(b - (hi + lo) / 2.0) <= 1e-12 * a
    oj.energy_rescale, oj.energy_shift = a, -b
    orp, oci, ov = og.assemble(oj)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    tol = 1e-12 * np.abs(ov) + 1e-14 * float(np.max(np.abs(ov)))
    assert np.all(np.abs(v - ov) <= tol)

    # FP64 recursion against the oracle's faithful KPM on 16 of the vectors
    n = geom.n
    mu64 = dev.kpm(M, random_phase=(R, 0), scheme=L.SCHEME_DOUBLING)
    print("kernel", dev.timing().kernel, flush=True)
    vecs = oracle_mod.random_phase(n, 16, 0)
    omu = oracle_mod.kpm(orp, oci, ov, M, vecs=vecs, doubling_all=True, nthreads=16)
    dmu = float(np.max(np.abs(mu64[:16] - omu)))
    print("FP64 max|dmu|/mu0 = %.3e" % (dmu / n), flush=True)
    assert dmu <= 1e-12 * n

    # mixed precision against the FP64 run
    mu32 = dev.kpm(M, random_phase=(R, 0), scheme=L.SCHEME_DOUBLING, precision=L.PREC_MIXED)
    assert np.max(np.abs(mu32 - mu64)) <= 2e-5 * n
    E, r64 = dev.dos(mu64.mean(axis=0) / n, a, b)
    _, r32 = dev.dos(mu32.mean(axis=0) / n, a, b)
    r64, r32 = r64[0], r32[0]
    x = (E - b) / a
    inner = np.abs(x) <= 0.99
    dd = np.abs(r32 - r64)
    print("mixed max|dDoS| = %.3e states/eV/atom (all bins, at x = %.6f), %.3e (|x| <= 0.99)"
          % (dd.max(), x[int(np.argmax(dd))], dd[inner].max()), flush=True)
    assert dd[inner].max() <= 1e-6
    assert dd.max() <= 3e-4


def test_c1_full_cell_mixed_dos(dev):
    """The headline cell itself, (32, 31) tiled 8 x 8: the FP32 recursion's DoS
    within 1e-6 states/eV/atom of the FP64 run at every one of the 2048
    energies (moments within 2e-5 mu_0)."""
    text = W.tblg_cell_deck(32, 31, 8, M=M)
    geom = L.Geometry(L.Deck(text))
    assert geom.n == 762112
    dev.set_geometry(geom)
    dev.pairs_build()
    job, a, b, nnz = W.assemble_rescaled(dev, geom, K_FRAC)
    n = geom.n
    mu64 = dev.kpm(M, random_phase=(R, 0), scheme=L.SCHEME_DOUBLING)
    mu32 = dev.kpm(M, random_phase=(R, 0), scheme=L.SCHEME_DOUBLING, precision=L.PREC_MIXED)
    assert np.max(np.abs(mu32 - mu64)) <= 2e-5 * n
    _, r64 = dev.dos(mu64.mean(axis=0) / n, a, b)
    _, r32 = dev.dos(mu32.mean(axis=0) / n, a, b)
    ddos = float(np.max(np.abs(r32 - r64)))
    print("8x8 mixed max|dDoS| = %.3e states/eV/atom" % ddos, flush=True)
    assert ddos <= 1e-6
