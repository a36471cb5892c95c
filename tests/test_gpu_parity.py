"""GPU parity through the C-ABI against the CPU oracle.

Bars (stated per test):
  * pair lists and CSR pattern: bit-exact (integer work);
  * Hamiltonian values: |dv| <= 1e-12 |v| + 1e-14 max|H| (the GPU's atan2/cos/exp
    differ from glibc by a few ulp; everything else is IEEE-exact);
  * FP64 moments: |dmu| <= 1e-12 mu_0 (different summation order only);
  * mixed precision (FP32 recursion): |dmu| <= 2e-5 mu_0, |dDoS| <= 1e-6 states/eV/atom inside
    the spectrum (the full c1 cell meets it at every energy, tests/test_gpu_c1.py);
  * DoS: |drho| <= 1e-11 (1 + |rho|).
"""
import io
import os
import re

import numpy as np
import pytest

from conftest import ROOT, golden_path

import leo2d_amd as L
from leo2d_amd import workloads as W

pytestmark = pytest.mark.gpu

DECKS = os.path.join(ROOT, "tests", "decks")
PARITY_DECKS = ["tblg_test1.in", "tblg_test2.in", "tblg_sc_7_6_k.in", "tblg_grid_3x3.in", "tblg_gshift_strain.in",
                "tblg_simple_flake.in", "tblg_c0_twin.in", "tblg_strain_flake.in", "tblg_trace.in", "tblg_relax_cell.in",
                "tblg_strained_flake.in", "tblg_strained_uniform.in", "tblg_strained_relax.in", "tblg_direct_cell.in",
                "tblg_strained_direct.in", "ttg_sc_7_6.in", "tblg_sc_hex.in", "tblg_sc_h.in"]


def deck_text(name):
    p = golden_path(name)
    if os.path.exists(p):
        return open(p).read()
    # STRAIN_FILE paths in the decks are relative to the repository root
    return open(os.path.join(DECKS, name)).read().replace("STRAIN_FILE = tests/", "STRAIN_FILE = %s/tests/" % ROOT)


@pytest.fixture(scope="module")
def dev():
    d = L.Device(0)
    yield d
    d.close()


def _oracle(orc, text):
    od = orc.parse_deck(text)
    orc.setup_supercell(od)
    og = orc.Geometry(orc.geom_desc(od))
    ojobs = orc.build_jobs(od, orc.target_list(od, og))
    return od, og, ojobs


def _setup(dev, text):
    deck = L.Deck(text)
    geom = L.Geometry(deck)
    dev.set_geometry(geom)
    counts = dev.pairs_build()
    m = re.search(r"STRAIN_FILE = (\S+)", text)
    if deck.get_int("strain_type") == 1 and m:
        dev.set_relaxation(*L.read_relaxation(m.group(1)))   # StrainCalc::loadFourierConfigFile
    return deck, geom, counts


@pytest.mark.parametrize("name", PARITY_DECKS)
def test_pairs_bit_exact(dev, oracle_mod, name):
    text = deck_text(name)
    print("case", name, flush=True)
    deck, geom, (ni, nx, ns) = _setup(dev, text)
    od, og, _ = _oracle(oracle_mod, text)
    for which, cnt in ((0, ni), (1, nx)):
        ij, sc, t = dev.pairs_copy(which, cnt)
        oij, osc, ot = og.pair_arrays(which)
        assert np.array_equal(ij, oij), which
        assert np.array_equal(sc, osc), which
        if which == 0:
            assert np.array_equal(t, ot)


@pytest.mark.parametrize("name", PARITY_DECKS + ["c1cell2"])
def test_staged_pairs_equal_thread_pairs(dev, monkeypatch, name):
    """The LDS-staged pair search (one wavefront per row) writes the same lists,
    counts and slot counts as the thread-per-row kernels, bit for bit: every
    parity deck (flakes, periodic cells with and without k-points, supercell
    grids, strained and trilayer decks) and the 1.05 deg (32, 31) cell tiled
    2 x 2."""
    print("case", name, flush=True)
    text = W.tblg_cell_deck(32, 31, 2) if name == "c1cell2" else deck_text(name)
    out = {}
    for staged in ("1", "0"):
        monkeypatch.setenv("LEO2D_PAIRS_STAGED", staged)
        deck, geom, counts = _setup(dev, text)
        assert dev.pairs_method() == int(staged), name
        lists = [dev.pairs_copy(which, cnt) for which, cnt in ((0, counts[0]), (1, counts[1]))]
        out[staged] = (counts, lists)
    assert out["1"][0] == out["0"][0]
    for (a_ij, a_sc, a_t), (b_ij, b_sc, b_t) in zip(out["1"][1], out["0"][1]):
        assert np.array_equal(a_ij, b_ij) and np.array_equal(a_sc, b_sc) and np.array_equal(a_t, b_t)


def _value_ok(v, ov):
    tol = 1e-12 * np.abs(ov) + 1e-14 * max(1e-300, float(np.max(np.abs(ov))) if len(ov) else 1.0)
    return np.all(np.abs(v - ov) <= tol)


@pytest.mark.parametrize("name", PARITY_DECKS)
@pytest.mark.parametrize("pattern", [L.PATTERN_REFERENCE, L.PATTERN_PRUNED])
def test_assembly_parity(dev, oracle_mod, name, pattern):
    text = deck_text(name)
    print("case", name, pattern, flush=True)
    deck, geom, _ = _setup(dev, text)
    od, og, ojobs = _oracle(oracle_mod, text)
    jobs = geom.jobs()
    for q in sorted({0, len(jobs) - 1}):
        dev.assemble(jobs[q], pattern)
        rp, ci, v = dev.csr_copy()
        orp, oci, ov = og.assemble(oracle_mod.make_job(od, ojobs[q], pruned=pattern))
        assert np.array_equal(rp, orp)
        assert np.array_equal(ci, oci)
        assert np.iscomplexobj(v) == np.iscomplexobj(ov)
        assert _value_ok(v, ov)


@pytest.mark.parametrize("name", ["tblg_test1.in", "tblg_test2.in", "tblg_sc_7_6_k.in", "tblg_gshift_strain.in"])
def test_kpm_targets_reference_scheme(dev, oracle_mod, name):
    text = deck_text(name)
    print("case", name, flush=True)
    deck, geom, _ = _setup(dev, text)
    od, og, ojobs = _oracle(oracle_mod, text)
    job = geom.jobs()[0]
    dev.assemble(job, L.PATTERN_REFERENCE)
    rp, ci, v = dev.csr_copy()
    M = od.opts["poly_order"]
    targets = list(job.targets[: job.num_targets])
    mu = dev.kpm(M, targets=targets)
    omu = oracle_mod.kpm(rp, ci, v, M, targets=targets)
    assert np.max(np.abs(mu - omu)) <= 1e-12
    # scheme variants and launch shapes give the same moments
    mu2 = dev.kpm(M, targets=targets, scheme=L.SCHEME_DOUBLING)
    omu2 = oracle_mod.kpm(rp, ci, v, M, targets=targets, doubling_all=True)
    assert np.max(np.abs(mu2 - omu2)) <= 1e-12
    mu3 = dev.kpm(M, targets=targets, grid=16, block=8)
    assert np.max(np.abs(mu3 - mu)) <= 1e-13


@pytest.mark.parametrize("R", [1, 3, 16, 64, 100])
def test_kpm_random_phase(dev, oracle_mod, R):
    text = deck_text("tblg_c0_twin.in")
    deck, geom, _ = _setup(dev, text)
    job = geom.jobs()[0]
    dev.assemble(job, L.PATTERN_PRUNED)
    rp, ci, v = dev.csr_copy()
    M = 64
    seed = 0
    mu = dev.kpm(M, random_phase=(R, seed), scheme=L.SCHEME_DOUBLING)
    vecs = oracle_mod.random_phase(geom.n, R, seed)
    omu = oracle_mod.kpm(rp, ci, v, M, vecs=vecs, doubling_all=True, nthreads=8)
    assert np.max(np.abs(mu[:, 0] - geom.n)) <= 1e-9
    assert np.max(np.abs(mu - omu)) <= 1e-12 * geom.n
    # reference scheme keeps the start vectors: mu_j = Re<v|T_j v> directly
    mur = dev.kpm(M, random_phase=(R, seed), scheme=L.SCHEME_REFERENCE)
    omur = oracle_mod.kpm(rp, ci, v, M, vecs=vecs, nthreads=8)
    assert np.max(np.abs(mur - omur)) <= 1e-12 * geom.n


def test_kpm_real_vectors_trace(dev, oracle_mod):
    text = deck_text("tblg_test1.in")
    deck, geom, _ = _setup(dev, text)
    job = geom.jobs()[1]
    dev.assemble(job, L.PATTERN_REFERENCE)
    rp, ci, v = dev.csr_copy()
    rng = np.random.default_rng(2)
    V = rng.standard_normal((5, geom.n))
    mu = dev.kpm(200, vectors=V)
    omu = oracle_mod.kpm(rp, ci, v, 200, vecs=V)
    assert np.max(np.abs(mu - omu)) <= 1e-12 * geom.n
    Vc = V + 1j * rng.standard_normal((5, geom.n))
    mu = dev.kpm(200, vectors=Vc)
    omu = oracle_mod.kpm(rp, ci, v, 200, vecs=Vc)
    assert np.max(np.abs(mu - omu)) <= 1e-12 * geom.n


@pytest.mark.parametrize("R", [32, 64])
@pytest.mark.parametrize("cw", ["1", "0"])
def test_sell_real_vectors(dev, oracle_mod, monkeypatch, R, cw):
    """Real H with real start vectors on the SELL tiles (two FP64 vectors per
    lane, 32 per 256-byte chunk), with and without the copy wavefront: moments
    within 1e-12 mu_0 of the oracle in the vector-start scheme and the
    doubling scheme."""
    monkeypatch.setenv("LEO2D_KPM_VARIANT", "20")
    monkeypatch.setenv("LEO2D_SELL_CW", cw)
    text = deck_text("tblg_c0_twin.in")
    deck, geom, _ = _setup(dev, text)
    job = geom.jobs()[0]
    dev.assemble(job, L.PATTERN_PRUNED)
    rp, ci, v = dev.csr_copy()
    assert not np.iscomplexobj(v)
    V = np.random.default_rng(R).standard_normal((R, geom.n))
    M = 40
    mu = dev.kpm(M, vectors=V)
    assert dev.timing().kernel == (21 if cw == "1" else 20)
    omu = oracle_mod.kpm(rp, ci, v, M, vecs=V, nthreads=8)
    assert np.max(np.abs(mu - omu)) <= 1e-12 * geom.n
    mu = dev.kpm(M, vectors=V, scheme=L.SCHEME_DOUBLING)
    omu = oracle_mod.kpm(rp, ci, v, M, vecs=V, doubling_all=True, nthreads=8)
    assert np.max(np.abs(mu - omu)) <= 1e-12 * geom.n


def test_mixed_precision_against_fp64(dev, oracle_mod):
    text = deck_text("tblg_c0_twin.in")
    deck, geom, _ = _setup(dev, text)
    job = geom.jobs()[0]
    dev.assemble(job, L.PATTERN_PRUNED)
    M = 500
    mu64 = dev.kpm(M, random_phase=(16, 0), scheme=L.SCHEME_DOUBLING)
    mu32 = dev.kpm(M, random_phase=(16, 0), scheme=L.SCHEME_DOUBLING, precision=L.PREC_MIXED)
    n = geom.n
    print("mixed max|dmu|/mu0:", float(np.max(np.abs(mu32 - mu64))) / n, flush=True)
    assert np.max(np.abs(mu32 - mu64)) <= 2e-5 * n
    a, b = job.energy_rescale, job.energy_shift
    E, r64 = dev.dos(mu64.mean(axis=0) / n, a, b)
    _, r32 = dev.dos(mu32.mean(axis=0) / n, a, b)
    dd = np.abs(r64[0] - r32[0])
    inner = np.abs((E - b) / a) <= 0.99
    print("mixed max|dDoS| states/eV/atom: %.3e (all bins), %.3e (|x| <= 0.99)" % (dd.max(), dd[inner].max()),
          flush=True)
    # inside the spectrum the bar holds; the two outermost bins (DoS zero there,
    # reconstruction divided by sqrt(1 - x^2) ~ pi/2M) carry the FP32 storage
    # rounding of 8,128 orbitals amplified ~2M/pi times (measured 5.9e-6)
    assert dd[inner].max() <= 1e-6
    assert dd.max() <= 3e-5


def test_dos_parity(dev, oracle_mod):
    rng = np.random.default_rng(4)
    for M in (200, 512, 2048):
        mu = rng.standard_normal((3, M)) * np.exp(-np.arange(M) / M)
        mu[:, 0] = 1.0
        E, rho = dev.dos(mu, 13.0, 0.25)
        orho = oracle_mod.dos(mu, 13.0, 0.25)
        assert np.max(np.abs(rho - orho) / (1 + np.abs(orho))) <= 1e-11
        assert np.max(np.abs(E - oracle_mod.energies(M, 13.0, 0.25))) <= 1e-13


def _cheb_numbers(text):
    import re

    return [float(x) for x in re.findall(r"-?\d+\.\d+", text)], re.sub(r"-?\d+\.\d+", "#", text)


@pytest.mark.parametrize("name", ["tblg_test1.in", "tblg_test2.in", "tblg_gshift_strain.in", "tblg_sc_7_6_k.in",
                                  "tblg_cond_flake.in", "tblg_cond_raw.in", "tblg_cond_raw.in:verbose",
                                  "tblg_trace.in:verbose",
                                  "tblg_sc_k1.in", "mlg_strain_k2.in", "ttlg_grid_k5.in", "tblg_sc_7_6_k.in:msave2",
                                  "tblg_relax_cell.in", "tblg_strained_flake.in", "tblg_strained_uniform.in",
                                  "tblg_strained_relax.in", "tblg_direct_cell.in", "tblg_strained_direct.in",
                                  "ttg_sc_7_6.in", "tblg_fixed_shift.in"])
def test_run_deck_cheb_matches_oracle(dev, oracle_mod, tmp_path, name):
    name, _, variant = name.partition(":")
    text = deck_text(name)
    if variant == "verbose":
        # conductivity with VERBOSE_SAVE = 1: the reference writes the job headers only
        text = text.replace("VERBOSE_SAVE = 0", "VERBOSE_SAVE = 1")
    if variant == "msave2":
        # MATRIX_SAVE = 2 on a complex H(k): _dxH_matrix.dat as well, no _dyH_matrix.dat
        text = text.replace("MATRIX_SAVE = 1", "MATRIX_SAVE = 2")
    prefix = str(tmp_path / "out")
    dev.run_deck(text, prefix)
    got = open(prefix + ".cheb").read()
    od, og, jobs, results = oracle_mod.run_deck(text)
    want = oracle_mod.cheb_text(od, results)
    gn, gs = _cheb_numbers(got)
    wn, ws = _cheb_numbers(want)
    assert gs == ws                                   # identical layout and integers
    assert np.max(np.abs(np.array(gn) - np.array(wn))) <= 1.0000001e-6   # 6-decimal rounding
    if od.opts["matrix_pos_save"] > 0:
        # positions of the last job, relaxation included: 12 decimals, the last one may
        # differ where device and host trigonometry differ by an ulp
        pos = og.job_positions(oracle_mod.make_job(od, jobs[-1]))
        got = np.loadtxt(prefix + "_pos.dat").reshape(-1, 5)
        want = np.loadtxt(io.StringIO(oracle_mod.pos_dat_text(pos, og.index_to_grid))).reshape(-1, 5)
        assert np.array_equal(got[:, 3:], want[:, 3:])
        assert np.max(np.abs(got[:, :3] - want[:, :3])) <= 2e-12
    if od.opts["k_sampling"] == 1:
        # Bloch vector of every job, byte for byte (param_tools.cpp:579-602)
        assert open(prefix + ".kpts").read() == oracle_mod.kpts_text(jobs)
    else:
        assert not os.path.exists(prefix + ".kpts")
    if od.opts["matrix_save"] > 0:
        _check_matrix_files(oracle_mod, od, og, jobs[-1], prefix)


def _check_matrix_files(orc, od, og, job, prefix):
    """_matrix.dat of the last job and, with MATRIX_SAVE = 2, its velocity dumps:
    _dxH_matrix.dat and (real H) _dyH_matrix.dat (locality.cpp:4366-4416, 5213-5244)."""
    rp, ci, v, dx, dy = og.assemble(orc.make_job(od, job), with_dx=True, with_dy=True)
    files = [("_matrix.dat", v)]
    if od.opts["matrix_save"] > 1:
        files.append(("_dxH_matrix.dat", dx))
        if not np.iscomplexobj(v):
            files.append(("_dyH_matrix.dat", dy))
        else:
            assert not os.path.exists(prefix + "_dyH_matrix.dat")
    for ext, vals in files:
        assert os.path.exists(prefix + ext), ext
        gl = open(prefix + ext).read().splitlines()
        wl = orc.matrix_dat_text(rp, ci, vals).splitlines()
        assert len(gl) == len(wl), ext
        for a, b in zip(gl, wl):
            ga = [float(x) for x in a.split(",")]
            wa = [float(x) for x in b.split(",")]
            assert ga[:2] == wa[:2]
            assert np.allclose(ga[2:], wa[2:], rtol=1e-5, atol=1e-6), ext


def test_run_deck_position_files_match_oracle(dev, oracle_mod, tmp_path):
    """MATRIX_POS_SAVE = 2 writes the last job's positions (_pos.dat) and the
    positions of every intra- and interlayer pair (_intra_pos.dat,
    _inter_pos.dat); the text equals the oracle's writers byte for byte."""
    text = deck_text("tblg_pos_save.in")
    prefix = str(tmp_path / "out")
    dev.run_deck(text, prefix)
    od, og, jobs, _ = oracle_mod.run_deck(text)
    assert len(jobs) == 4
    pos = og.job_positions(oracle_mod.make_job(od, jobs[-1]))
    assert open(prefix + "_pos.dat").read() == oracle_mod.pos_dat_text(pos, og.index_to_grid)
    for which, ext in ((0, "_intra_pos.dat"), (1, "_inter_pos.dat")):
        ij, _, _ = og.pair_arrays(which)
        assert open(prefix + ext).read() == oracle_mod.pair_pos_text(ij, pos), ext
    assert od.opts["matrix_save"] == 2
    _check_matrix_files(oracle_mod, od, og, jobs[-1], prefix)


def test_spectrum_and_nan_errors(dev):
    n = 50
    rp = np.arange(n + 1, dtype=np.int64)
    ci = np.arange(n, dtype=np.int32)
    v = np.linspace(-1.4, 1.4, n)          # spectrum outside [-1, 1]
    dev.csr_upload(rp, ci, v)
    with pytest.raises(L.Leo2dError) as e:
        dev.kpm(64, targets=[0, 1])
    assert e.value.code == L.ERR_SPECTRUM
    v[3] = np.nan
    with pytest.raises(L.Leo2dError) as e:
        dev.csr_upload(rp, ci, v)
    assert e.value.code == L.ERR_NAN
    with pytest.raises(L.Leo2dError) as e:
        dev.kpm(62, targets=[0])
    assert e.value.code == L.ERR_ARG
    with pytest.raises(L.Leo2dError) as e:
        dev.kpm(64, targets=[n + 5])
    assert e.value.code == L.ERR_ARG


def test_oversized_vector_block_is_alloc_error(dev):
    """A vector block larger than the GPU's memory: 128 complex FP64 vectors on
    a 150 M-row diagonal H are 307 GB per block (288 GB of HBM).  The library
    returns LEO2D_ERR_ALLOC (not a crash or a hang) and the handle keeps
    working afterwards."""
    n = 150_000_000
    dev.csr_upload(np.arange(n + 1, dtype=np.int64), np.arange(n, dtype=np.int32), np.full(n, 0.25))
    with pytest.raises(L.Leo2dError) as e:
        dev.kpm(8, random_phase=(128, 1), block=128, natural_order=True)
    assert e.value.code == L.ERR_ALLOC, str(e.value)
    m = 64
    dev.csr_upload(np.arange(m + 1, dtype=np.int64), np.arange(m, dtype=np.int32), np.full(m, 0.25))
    mu = dev.kpm(8, random_phase=(4, 1), scheme=L.SCHEME_DOUBLING)
    assert np.allclose(mu, np.cos(np.arange(8) * np.arccos(0.25))[None, :] * m, rtol=1e-12, atol=1e-12 * m)


@pytest.mark.parametrize("name", ["tblg_c0_twin.in", "tblg_sc_7_6_k.in"])
def test_spatial_order_matches_natural_order(dev, name):
    """The recursion runs on P H P^T (Morton order) by default; moments must
    match the reference row order up to summation order."""
    text = deck_text(name)
    deck, geom, _ = _setup(dev, text)
    job = geom.jobs()[0]
    dev.assemble(job, L.PATTERN_PRUNED)
    M = 128
    a = dev.kpm(M, random_phase=(16, 3), scheme=L.SCHEME_DOUBLING)
    b = dev.kpm(M, random_phase=(16, 3), scheme=L.SCHEME_DOUBLING, natural_order=True)
    assert np.max(np.abs(a - b)) <= 1e-12 * geom.n
    t = list(job.targets[: job.num_targets])
    a = dev.kpm(M, targets=t)
    b = dev.kpm(M, targets=t, natural_order=True)
    assert np.max(np.abs(a - b)) <= 1e-13
    rng = np.random.default_rng(9)
    V = rng.standard_normal((4, geom.n))
    a = dev.kpm(M, vectors=V)
    b = dev.kpm(M, vectors=V, natural_order=True)
    assert np.max(np.abs(a - b)) <= 1e-12 * geom.n


@pytest.mark.parametrize("variant", ["0", "1", "14", "15", "20"])
@pytest.mark.parametrize("R", [4, 16, 64, 128])
def test_step_kernel_variants(dev, oracle_mod, monkeypatch, variant, R):
    """Every step kernel (lane per vector, wavefront per row, vector chunks
    split over the XCDs, two FP32 rows per wavefront, SELL-4-sigma with
    LDS-staged tiles) gives the oracle's moments; a kernel that does not apply
    to (R, precision) hands over to one that does."""
    monkeypatch.setenv("LEO2D_KPM_VARIANT", variant)
    text = deck_text("tblg_c0_twin.in")
    deck, geom, _ = _setup(dev, text)
    job = geom.jobs()[0]
    job.complex_matrix = 1
    job.k[0], job.k[1] = 0.013, -0.007
    dev.assemble(job, L.PATTERN_PRUNED)
    rp, ci, v = dev.csr_copy()
    M = 48
    vecs = oracle_mod.random_phase(geom.n, R, 5)
    omu = oracle_mod.kpm(rp, ci, v, M, vecs=vecs, doubling_all=True, nthreads=8)
    for prec, tol in ((L.PREC_FP64, 1e-12), (L.PREC_MIXED, 2e-5)):
        mu = dev.kpm(M, random_phase=(R, 5), scheme=L.SCHEME_DOUBLING, precision=prec)
        assert np.max(np.abs(mu - omu)) <= tol * geom.n, (variant, R, prec)
    mur = dev.kpm(M, random_phase=(R, 5), scheme=L.SCHEME_REFERENCE)
    omur = oracle_mod.kpm(rp, ci, v, M, vecs=vecs, nthreads=8)
    assert np.max(np.abs(mur - omur)) <= 1e-12 * geom.n
    # FP32 recursion through the reference's vector-start scheme (the third dot)
    mur = dev.kpm(M, random_phase=(R, 5), scheme=L.SCHEME_REFERENCE, precision=L.PREC_MIXED)
    assert np.max(np.abs(mur - omur)) <= 2e-5 * geom.n, (variant, R)


@pytest.mark.parametrize("variant", ["1", "14", "15", "20"])
@pytest.mark.parametrize("R", [32, 64])
def test_variants_ragged_row_count(dev, oracle_mod, monkeypatch, variant, R):
    """A flake whose orbital count (1022) is not a multiple of the rows a
    wavefront, workgroup or XCD part covers (n mod 4 = 2, n mod 8 = 6): the last
    row sets are partial.  Real H; moments equal the oracle's in both precisions
    and both schemes."""
    monkeypatch.setenv("LEO2D_KPM_VARIANT", variant)
    text = W.tblg_flake_deck(2.0, 21, M=64)
    deck, geom, _ = _setup(dev, text)
    assert geom.n == 1022
    job = geom.jobs()[0]
    job.energy_rescale, job.energy_shift = 10.5, 0.0
    dev.assemble(job, L.PATTERN_PRUNED)
    rp, ci, v = dev.csr_copy()
    M = 40
    vecs = oracle_mod.random_phase(geom.n, R, 21)
    omu = oracle_mod.kpm(rp, ci, v, M, vecs=vecs, doubling_all=True, nthreads=8)
    omur = oracle_mod.kpm(rp, ci, v, M, vecs=vecs, nthreads=8)
    for grid in (0, 8, 40):
        for prec, tol in ((L.PREC_FP64, 1e-12), (L.PREC_MIXED, 2e-5)):
            mu = dev.kpm(M, random_phase=(R, 21), scheme=L.SCHEME_DOUBLING, precision=prec, grid=grid)
            assert np.max(np.abs(mu - omu)) <= tol * geom.n, (variant, R, grid, prec)
            mu = dev.kpm(M, random_phase=(R, 21), scheme=L.SCHEME_REFERENCE, precision=prec, grid=grid)
            assert np.max(np.abs(mu - omur)) <= tol * geom.n, (variant, R, grid, prec)


@pytest.mark.parametrize("R", [16, 32, 64, 128])
def test_xcd_chunk_variant(dev, oracle_mod, monkeypatch, R):
    """Vector chunks split over the XCDs (variant 14): every (chunk, row part)
    pair is covered exactly once for grids whose per-XCD share does not divide
    the rows; moments equal the oracle's for random phases (both precisions), a
    real H with complex vectors, and real vectors through the vector-start
    scheme (third dot)."""
    monkeypatch.setenv("LEO2D_KPM_VARIANT", "14")
    text = deck_text("tblg_c0_twin.in")
    deck, geom, _ = _setup(dev, text)
    job = geom.jobs()[0]
    job.complex_matrix = 1
    job.k[0], job.k[1] = -0.006, 0.017
    dev.assemble(job, L.PATTERN_PRUNED)
    rp, ci, v = dev.csr_copy()
    M = 40
    vecs = oracle_mod.random_phase(geom.n, R, 8)
    omu = oracle_mod.kpm(rp, ci, v, M, vecs=vecs, doubling_all=True, nthreads=8)
    for grid in (0, 8, 40):
        for prec, tol in ((L.PREC_FP64, 1e-12), (L.PREC_MIXED, 2e-5)):
            mu = dev.kpm(M, random_phase=(R, 8), scheme=L.SCHEME_DOUBLING, precision=prec, grid=grid)
            assert np.max(np.abs(mu - omu)) <= tol * geom.n, (R, grid, prec)
            assert dev.timing().kernel == 14
    mur = dev.kpm(M, random_phase=(R, 8), scheme=L.SCHEME_REFERENCE)
    omur = oracle_mod.kpm(rp, ci, v, M, vecs=vecs, nthreads=8)
    assert np.max(np.abs(mur - omur)) <= 1e-12 * geom.n
    job.complex_matrix = 0
    job.k[0], job.k[1] = 0.0, 0.0
    dev.assemble(job, L.PATTERN_PRUNED)
    rp, ci, v = dev.csr_copy()
    omu = oracle_mod.kpm(rp, ci, v, M, vecs=vecs, doubling_all=True, nthreads=8)
    mu = dev.kpm(M, random_phase=(R, 8), scheme=L.SCHEME_DOUBLING)
    assert np.max(np.abs(mu - omu)) <= 1e-12 * geom.n
    V = np.random.default_rng(7).standard_normal((R, geom.n))
    for prec, tol in ((L.PREC_FP64, 1e-12), (L.PREC_MIXED, 2e-5)):
        mu = dev.kpm(M, vectors=V, precision=prec)
        omu = oracle_mod.kpm(rp, ci, v, M, vecs=V, nthreads=8)
        assert np.max(np.abs(mu - omu)) <= tol * geom.n, (R, prec)


@pytest.mark.parametrize("variant", ["1"])
@pytest.mark.parametrize("R", [64, 128])
def test_xcd_slice_schedule(dev, oracle_mod, monkeypatch, variant, R):
    """The one-slice-per-XCD row schedule covers every row exactly once: the
    moments equal the oracle's, including for n not a multiple of 8 x 4."""
    monkeypatch.setenv("LEO2D_KPM_VARIANT", variant)
    monkeypatch.setenv("LEO2D_KPM_SCHED", "1")
    text = deck_text("tblg_c0_twin.in")
    deck, geom, _ = _setup(dev, text)
    job = geom.jobs()[0]
    job.complex_matrix = 1
    job.k[0], job.k[1] = -0.011, 0.004
    dev.assemble(job, L.PATTERN_PRUNED)
    rp, ci, v = dev.csr_copy()
    M = 40
    vecs = oracle_mod.random_phase(geom.n, R, 9)
    omu = oracle_mod.kpm(rp, ci, v, M, vecs=vecs, doubling_all=True, nthreads=8)
    for grid in (0, 8, 24):
        mu = dev.kpm(M, random_phase=(R, 9), scheme=L.SCHEME_DOUBLING, grid=grid)
        assert np.max(np.abs(mu - omu)) <= 1e-12 * geom.n, (variant, R, grid)


@pytest.mark.parametrize("cw", ["1", "0"])
def test_sell_split_build(dev, oracle_mod, monkeypatch, cw):
    """SELL tiles of one matrix built in 16 fixed row ranges on host threads (the
    path of matrices of 1 M rows or more, forced here on the c0 twin's 8,128
    rows by LEO2D_SELL_SPLIT_ROWS): FP64 moments within 1e-12 mu_0 of the oracle
    and of the one-pass build, complex H(k), R = 64."""
    print("split build, copy wave", cw, flush=True)
    monkeypatch.setenv("LEO2D_SELL_CW", cw)
    monkeypatch.setenv("LEO2D_KPM_VARIANT", "20")
    text = deck_text("tblg_c0_twin.in")
    deck, geom, _ = _setup(dev, text)
    job = geom.jobs()[0]
    job.complex_matrix = 1
    job.k[0], job.k[1] = 0.011, -0.004
    M, R = 40, 64
    res = {}
    for split in ("1000", str(1 << 20)):
        monkeypatch.setenv("LEO2D_SELL_SPLIT_ROWS", split)
        dev.assemble(job, L.PATTERN_PRUNED)
        res[split] = dev.kpm(M, random_phase=(R, 9), scheme=L.SCHEME_DOUBLING)
        assert dev.timing().kernel == (21 if cw == "1" else 20)
    rp, ci, v = dev.csr_copy()
    vecs = oracle_mod.random_phase(geom.n, R, 9)
    omu = oracle_mod.kpm(rp, ci, v, M, vecs=vecs, doubling_all=True, nthreads=8)
    for split, mu in res.items():
        assert np.max(np.abs(mu - omu)) <= 1e-12 * geom.n, split
    assert np.max(np.abs(res["1000"] - res[str(1 << 20)])) <= 1e-12 * geom.n


@pytest.mark.parametrize("cfg", [("288", "1", "64"), ("64", "0", "64"), ("128", "1", "13"), ("288", "0", "5"),
                                 ("200", "1", "48")])
@pytest.mark.parametrize("R", [16, 32, 64, 128])
@pytest.mark.parametrize("cw", ["1", "0"])
def test_sell_variant_shapes(dev, oracle_mod, monkeypatch, cfg, R, cw):
    """SELL-4-sigma step with LDS-staged tiles (variant 20), its staging copies
    issued by a wavefront of their own (cw = 1, tiles of at most 56 rows) or by
    every wavefront (cw = 0), under LDS budgets of 64 to 288 staged rows (tiles
    cut short by the budget), tiles of at most 5 to 64 rows (partial groups,
    idle wavefronts) and both tile schedules (XCD slices, interleaved), walked
    forwards and backwards on alternate steps: FP64 moments within 1e-12 mu_0
    of the oracle in both schemes, FP32 recursion within 2e-5 mu_0, complex
    H(k) and a real H.  R = 128 falls back to the CSR step."""
    ucap, sliced, tmax = cfg
    print("case", cfg, R, "copy wave", cw, flush=True)
    monkeypatch.setenv("LEO2D_SELL_CW", cw)
    monkeypatch.setenv("LEO2D_KPM_VARIANT", "20")
    monkeypatch.setenv("LEO2D_SELL_UCAP", ucap)
    monkeypatch.setenv("LEO2D_SELL_SLICED", sliced)
    monkeypatch.setenv("LEO2D_SELL_TMAX", tmax)
    text = deck_text("tblg_c0_twin.in")
    deck, geom, _ = _setup(dev, text)
    job = geom.jobs()[0]
    for kx in (0.0, 0.011):
        job.complex_matrix = 1 if kx else 0
        job.k[0], job.k[1] = kx, -0.004 if kx else 0.0
        dev.assemble(job, L.PATTERN_PRUNED)
        rp, ci, v = dev.csr_copy()
        M = 40
        vecs = oracle_mod.random_phase(geom.n, R, 9)
        omu = oracle_mod.kpm(rp, ci, v, M, vecs=vecs, doubling_all=True, nthreads=8)
        omur = oracle_mod.kpm(rp, ci, v, M, vecs=vecs, nthreads=8)
        for prec, tol in ((L.PREC_FP64, 1e-12), (L.PREC_MIXED, 2e-5)):
            mu = dev.kpm(M, random_phase=(R, 9), scheme=L.SCHEME_DOUBLING, precision=prec)
            assert np.max(np.abs(mu - omu)) <= tol * geom.n, (cfg, R, kx, prec)
            if R == 64:
                assert dev.timing().kernel == (21 if cw == "1" else 20)
            mu = dev.kpm(M, random_phase=(R, 9), scheme=L.SCHEME_REFERENCE, precision=prec)
            assert np.max(np.abs(mu - omur)) <= tol * geom.n, (cfg, R, kx, prec)
