"""Pin the CPU oracle before trusting it.

* couplings and sheet pair lists against the reference's own graphene.cpp and
  sheet.cpp compiled into oracle/_ref (skipped when the reference tree is
  absent);
* SpMV against the reference's SpMatrix::vectorMultiply (oracle/_ref);
* closed forms: known coupling values, KPM moments from a dense
  eigendecomposition, DCT-III direct sum, Jackson g_0 = 1, normalisation;
* the reference's golden fixtures tests/golden/tblg_test{1,2}: target indices
  and the Chebyshev-Gauss normalisation of every golden DoS row.
"""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from conftest import golden_path


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- couplings
def test_intralayer_known_answers(oracle_mod):
    L = oracle_mod.lib()
    # graphene.cpp:63-65: onsite 0.3504, nearest neighbour -2.8922 (A row, B col at v=(0,0))
    assert L.orc_intralayer(0, 0, 0, 0, 0) == 0.3504
    assert L.orc_intralayer(0, 0, 1, 0, 0) == -2.8922
    assert L.orc_intralayer(0, 1, 0, 0, 0) == -2.8922
    assert L.orc_intralayer(0, 0, 0, 1, 0) == 0.2425          # r = 3: second shell (same sublattice)
    assert L.orc_intralayer(0, 0, 0, 2, 0) == -0.0209         # r = 12
    assert L.orc_intralayer(0, 0, 0, 5, 5) == 0.0             # far
    # simple model: only the nearest-neighbour shell
    assert L.orc_intralayer(1, 0, 1, 0, 0) == -2.7
    assert L.orc_intralayer(1, 0, 0, 0, 0) == 0.0
    assert L.orc_intralayer(1, 0, 0, 1, 0) == 0.0


def test_interlayer_closed_forms(oracle_mod):
    L = oracle_mod.lib()
    # r >= 8 A -> 0 and |z| > 5 (after the int truncation of graphene.cpp:117) -> 0
    assert L.orc_interlayer(0, 0, 0, 8.0, 0.0, 3.35, 0.0, 0.0) == 0.0
    assert L.orc_interlayer(0, 0, 0, 1.0, 0.0, 6.5, 0.0, 0.0) == 0.0
    # at z = 3.35, z_eps = 0: coefficients reduce to their constant terms
    r, th = 1.3, 0.4
    x, y = r * math.cos(th), r * math.sin(th)
    rs = r / 2.4768
    V0 = 0.310 * math.exp(-1.750 * rs * rs) * math.cos(1.990 * rs)
    V3 = -0.068 * rs * rs * math.exp(-3.286 * (rs - 0.5) ** 2)
    V6 = -0.008 * math.exp(-2.272 * (rs - 1.217) ** 2) * math.sin(1.562 * rs)
    t21 = th - 0.0 - math.pi / 6
    t12 = th - 0.0 + math.pi - math.pi / 6
    want = V0 + V3 * (math.cos(3 * t12) + math.cos(3 * t21)) + V6 * (math.cos(6 * t12) + math.cos(6 * t21))
    got = L.orc_interlayer(0, 0, 0, x, y, 3.35, 0.0, 0.0)
    assert abs(got - want) < 1e-14
    # smooth cutoff factor exp(1 - 1/d^2) at d = 0.5 is e^-3
    r = 7.5
    rs = r / 2.4768
    base_v0 = 0.310 * math.exp(-1.750 * rs * rs) * math.cos(1.990 * rs)
    v3 = -0.068 * rs * rs * math.exp(-3.286 * (rs - 0.5) ** 2)
    v6 = -0.008 * math.exp(-2.272 * (rs - 1.217) ** 2) * math.sin(1.562 * rs)
    t21 = 0.0 - math.pi / 6
    t12 = math.pi - math.pi / 6
    base = base_v0 + v3 * (math.cos(3 * t12) + math.cos(3 * t21)) + v6 * (math.cos(6 * t12) + math.cos(6 * t21))
    got = L.orc_interlayer(0, 0, 0, r, 0.0, 3.35, 0.0, 0.0)
    assert abs(got - base * math.exp(-3.0)) < 1e-18
    # simple model closed form
    d = math.sqrt(1.0 + 3.35 ** 2)
    assert abs(L.orc_interlayer(1, 0, 0, 1.0, 0.0, 3.35, 0, 0) - 0.48 * math.exp(-(d - 3.35) / 0.45)) < 1e-15
    assert L.orc_interlayer(1, 0, 0, 6.0, 0.0, 0.0, 0, 0) == 0.0


def test_couplings_match_reference_build(oracle_mod, ref_lib):
    L = oracle_mod.lib()
    for orow in range(2):
        for ocol in range(2):
            for v0 in range(-6, 7):
                for v1 in range(-6, 7):
                    assert L.orc_intralayer(0, orow, ocol, v0, v1) == ref_lib.ref_graphene_intra(orow, ocol, v0, v1)
    rng = np.random.default_rng(7)
    zs = [3.35, 3.4, -3.35, 5.5, -5.9, 6.0, 6.5, 0.0]
    for _ in range(4000):
        x, y = rng.uniform(-9, 9, 2)
        z = zs[rng.integers(len(zs))] if rng.random() < 0.5 else rng.uniform(-7, 7)
        th1, th2 = rng.uniform(-0.5, 0.5, 2)
        o1, o2 = rng.integers(0, 2, 2)
        a = L.orc_interlayer(0, int(o1), int(o2), x, y, z, th1, th2)
        b = ref_lib.ref_graphene_inter(int(o1), int(o2), x, y, z, th1, th2)
        assert a == b, (x, y, z, th1, th2, o1, o2, a, b)


def test_strained_graphene_couplings_match_reference_build(oracle_mod, ref_lib):
    """StrainedGraphene intralayer terms (strained_graphene.cpp:16, 300) from the
    reference build, bit for bit: the unstrained table over a window of hops, and
    the strained form over random hop directions, strains and sheet angles."""
    L = oracle_mod.lib()
    L.orc_strained_intra.restype = C.c_double
    L.orc_strained_intra.argtypes = [C.c_int] * 4 + [C.c_double, C.c_double, C.c_void_p, C.c_double]
    for orow in range(2):
        for ocol in range(2):
            for v0 in range(-7, 8):
                for v1 in range(-7, 8):
                    assert L.orc_intralayer(3, orow, ocol, v0, v1) == ref_lib.ref_strained_intra(orow, ocol, v0, v1)
    rng = np.random.default_rng(11)
    for _ in range(3000):
        orow, ocol = (int(x) for x in rng.integers(0, 2, 2))
        v0, v1 = (int(x) for x in rng.integers(-3, 4, 2))
        t = rng.uniform(-math.pi, math.pi)
        h = np.array([math.cos(t), math.sin(t)])
        u = rng.normal(0, 0.05, 4)
        th = rng.uniform(-0.3, 0.3)
        a = L.orc_strained_intra(orow, ocol, v0, v1, h[0], h[1], _p(u), th)
        b = ref_lib.ref_strained_intra_hop(orow, ocol, v0, v1, h[0], h[1], _p(u), th)
        assert a == b, (orow, ocol, v0, v1, a, b)


# ---------------------------------------------------------------- sheet pairs
def _ref_sheet_pairs(ref_lib, min0, max0, max_R, bc, sc=None, st=None, S=1):
    sca = np.asarray(sc if sc is not None else [[0, 0], [0, 0]], dtype=np.float64).reshape(4)
    sta = np.asarray(st if st is not None else [[0, 0], [0, 0]], dtype=np.int32).reshape(4)
    gs = np.zeros(2)
    nor = C.c_int()
    np_ = ref_lib.ref_sheet_intra_pairs(min0, min0, max0, max0, max_R, bc, _p(sca), _p(sta), S, 0, _p(gs), 0,
                                        None, None, None, 0, C.byref(nor))
    ij = np.zeros(2 * np_, dtype=np.int32)
    scv = np.zeros(2 * np_, dtype=np.int32)
    t = np.zeros(np_)
    ref_lib.ref_sheet_intra_pairs(min0, min0, max0, max0, max_R, bc, _p(sca), _p(sta), S, 0, _p(gs), 0,
                                  _p(ij), _p(scv), _p(t), np_, C.byref(nor))
    pos = np.zeros(3 * nor.value)
    ref_lib.ref_sheet_positions(min0, min0, max0, max0, max_R, bc, _p(sca), _p(sta), _p(pos), nor.value)
    return nor.value, ij.reshape(-1, 2), scv.reshape(-1, 2), t, pos


def _oracle_single_sheet(orc, min0, max0, max_R, bc, sc=None, st=None, S=1, mat=0):
    d = orc.GeomDesc()
    d.num_sheets = 1
    d.boundary_condition = bc
    d.sc_search_size = S
    sh = d.sheet[0]
    sh.mat = mat
    sh.min_shape[0] = sh.min_shape[1] = min0
    sh.max_shape[0] = sh.max_shape[1] = max0
    sh.max_R = max_R
    sh.boundary_condition = bc
    if bc:
        for a in range(2):
            for b in range(2):
                sh.supercell[a][b] = sc[a][b]
                sh.supercell_stride[a][b] = st[a][b]
                d.supercell[a][b] = sc[a][b]
    return orc.Geometry(d)


@pytest.mark.parametrize("case", ["flake10", "flake14", "periodic_7_6_sheet0", "periodic_7_6_sheet1", "grid_3x3",
                                  "strained:flake10", "strained:periodic_7_6_sheet1"])
def test_sheet_pairs_match_reference_build(oracle_mod, ref_lib, case):
    """Enumeration, positions and intralayer pairs of one sheet against the
    reference's Sheet; 'strained:' cases build StrainedGraphene sheets (its own
    lattice and orbital constants, a 6-cell search window and its own table)."""
    orc = oracle_mod
    strained = case.startswith("strained:")
    case = case.partition(":")[2] if strained else case
    ref_lib.ref_set_sheet_material(1 if strained else 0)
    uc = orc.lattice(orc.MAT_STRAINED_GRAPHENE if strained else orc.MAT_GRAPHENE)
    if case.startswith("flake"):
        R = int(case[5:])
        args = dict(min0=-R, max0=R, max_R=float(R), bc=0)
    elif case == "grid_3x3":
        sc = [[3 * uc[0][0], 3 * uc[0][1]], [3 * uc[1][0], 3 * uc[1][1]]]
        args = dict(min0=-2, max0=6, max_R=6.0, bc=1, sc=sc, st=[[3, 0], [0, 3]])
    else:
        M, N = 7, 6
        A = ((N, M), (-M, M + N)) if case.endswith("0") else ((M, N), (-N, M + N))
        sc = [[A[r][0] * uc[0][0] + A[r][1] * uc[1][0], A[r][0] * uc[0][1] + A[r][1] * uc[1][1]] for r in range(2)]
        args = dict(min0=-8, max0=21, max_R=21.0, bc=1, sc=sc, st=[list(A[0]), list(A[1])])
    nref, ij, scv, t, pos = _ref_sheet_pairs(ref_lib, **args)
    ref_lib.ref_set_sheet_material(0)
    g = _oracle_single_sheet(orc, mat=orc.MAT_STRAINED_GRAPHENE if strained else 0, **args)
    assert g.n == nref
    i2g, i2p = g.index_to_grid, g.index_to_pos
    # single unrotated sheet at height 0: global positions are the sheet positions
    assert np.array_equal(i2p, pos)
    oij, osc, ot = g.pair_arrays(0)
    assert np.array_equal(oij, ij)
    assert np.array_equal(osc, scv)
    assert np.array_equal(ot, t)
    if case == "periodic_7_6_sheet0":
        assert nref == 2 * (7 * 7 + 7 * 6 + 6 * 6)  # one layer of the 4(N^2+NM+M^2) cell


def test_spmv_matches_reference_build(oracle_mod, ref_lib):
    orc = oracle_mod
    deck, geom, jobs, _ = orc.run_deck(open(golden_path("tblg_test2.in")).read())
    for cpx in (0, 1):
        job = orc.make_job(deck, jobs[1])
        job.complex_matrix = cpx
        rp, ci, v = geom.assemble(job)
        n = len(rp) - 1
        rng = np.random.default_rng(3)
        w = 2 if cpx else 1
        x = rng.standard_normal(w * n)
        y0 = rng.standard_normal(w * n)
        vv = np.ascontiguousarray(v.view(np.float64) if cpx else v)
        for alpha, beta in ((1.0, 0.0), (2.0, 0.0), (2.0, 1.0)):
            ya = y0.copy()
            yb = y0.copy()
            orc.lib().orc_spmv(n, _p(rp), _p(ci), _p(vv), cpx, _p(x), _p(ya), alpha, beta)
            rp32 = rp.astype(np.int32)
            ref_lib.ref_spmv(n, len(ci), _p(rp32), _p(ci), _p(vv), cpx, _p(x), _p(yb), alpha, beta)
            assert np.array_equal(ya, yb)


# ---------------------------------------------------------------- inter pairs
def test_inter_pairs_against_all_pairs_scan(oracle_mod):
    """Hstruct is not buildable here (needs fftw3.h): pin the window search by
    an all-pairs scan.  Every stored pair is within 12.5 A in-plane, every
    pair within the 8 A coupling range is stored, and rows are in order."""
    orc = oracle_mod
    deck, geom, jobs, _ = orc.run_deck(open(golden_path("tblg_test1.in")).read())
    ij, sc, _ = geom.pair_arrays(1)
    p = geom.index_to_pos.reshape(-1, 3)
    s = geom.index_to_grid.reshape(-1, 4)[:, 3]
    d2 = (p[ij[:, 0], 0] - p[ij[:, 1], 0]) ** 2 + (p[ij[:, 0], 1] - p[ij[:, 1], 1]) ** 2
    assert np.all(d2 < 12.5 ** 2)
    assert np.all(np.abs(s[ij[:, 0]] - s[ij[:, 1]]) == 1)
    assert np.all(np.diff(ij[:, 0]) >= 0)
    have = set(map(tuple, ij.tolist()))
    n = geom.n
    for a in range(n):
        for b in range(n):
            if abs(int(s[a]) - int(s[b])) != 1:
                continue
            if (p[a, 0] - p[b, 0]) ** 2 + (p[a, 1] - p[b, 1]) ** 2 < 8.0 ** 2:
                assert (a, b) in have, (a, b)
    # flake: no images
    assert np.all(sc == 0)


def test_strained_graphene_inter_window(oracle_mod):
    """StrainedGraphene searches 8 cells each way for interlayer pairs (int(8 + a/3),
    hstruct.cpp:305): on a flake that window holds every pair within the 12.5 A
    cut, so the stored list is exactly the all-pairs scan at 12.5 A (graphene's
    6-cell window misses some between 10 and 12.5 A)."""
    orc = oracle_mod
    deck, geom, jobs, _ = orc.run_deck(open(os.path.join(DECKS, "tblg_strained_flake.in")).read())
    assert all(sd.mat == orc.MAT_STRAINED_GRAPHENE for sd in deck.sdata)
    ij, _, _ = geom.pair_arrays(1)
    p = geom.index_to_pos.reshape(-1, 3)
    s = geom.index_to_grid.reshape(-1, 4)[:, 3]
    d2 = (p[:, None, 0] - p[None, :, 0]) ** 2 + (p[:, None, 1] - p[None, :, 1]) ** 2
    scan = np.argwhere((np.abs(s[:, None] - s[None, :]) == 1) & (d2 < 12.5 ** 2))
    assert sorted(map(tuple, ij.tolist())) == sorted(map(tuple, scan.tolist()))


# ---------------------------------------------------------------- KPM
def _dense(rp, ci, v):
    n = len(rp) - 1
    H = np.zeros((n, n), dtype=v.dtype)
    for i in range(n):
        for q in range(rp[i], rp[i + 1]):
            H[i, ci[q]] += v[q]
    return H


@pytest.mark.parametrize("deckname,cpx", [("tblg_test1.in", False), ("tblg_test2.in", True)])
def test_kpm_against_dense_eigendecomposition(oracle_mod, deckname, cpx):
    orc = oracle_mod
    deck, geom, jobs, _ = orc.run_deck(open(golden_path(deckname)).read())
    rp, ci, v = geom.assemble(orc.make_job(deck, jobs[0]))
    assert np.iscomplexobj(v) == cpx
    H = _dense(rp, ci, v)
    assert np.allclose(H, H.conj().T, atol=1e-15)     # Hermitian
    lam, U = np.linalg.eigh(H)
    assert np.max(np.abs(lam)) < 1.0
    M = 64
    targets = jobs[0]["target_list"]
    mu = orc.kpm(rp, ci, v, M, targets=targets)
    Tn = np.cos(np.outer(np.arange(M), np.arccos(lam)))   # T_n(lambda_k)
    for r, t in enumerate(targets):
        w = np.abs(U[t, :]) ** 2
        exact = Tn @ w
        assert abs(mu[r, 0] - 1.0) == 0.0
        assert abs(mu[r, 1] - H[t, t].real) < 1e-15
        assert np.max(np.abs(mu[r] - exact)) < 1e-12
    # vector starts: mu_n = <v|T_n(H)|v>
    rng = np.random.default_rng(5)
    V = rng.standard_normal((3, geom.n))
    muv = orc.kpm(rp, ci, v, M, vecs=V)
    for r in range(3):
        c = U.conj().T @ V[r]
        exact = Tn @ (np.abs(c) ** 2)
        assert np.max(np.abs(muv[r] - exact)) < 1e-10 * np.sum(V[r] ** 2)
    # the doubling-from-j=1 scheme agrees with the reference scheme
    mud = orc.kpm(rp, ci, v, M, vecs=V, doubling_all=True)
    assert np.max(np.abs(mud - muv)) < 1e-10 * geom.n


def test_kpm_threads_bitwise_equal(oracle_mod):
    orc = oracle_mod
    deck, geom, jobs, _ = orc.run_deck(open(golden_path("tblg_test1.in")).read())
    rp, ci, v = geom.assemble(orc.make_job(deck, jobs[0]))
    V = orc.random_phase(geom.n, 6, 11)
    a = orc.kpm(rp, ci, v, 40, vecs=V, nthreads=1)
    b = orc.kpm(rp, ci, v, 40, vecs=V, nthreads=4)
    assert np.array_equal(a, b)


def test_random_phase_hash_independent_restatement(oracle_mod):
    L = oracle_mod.lib()
    MASK = (1 << 64) - 1

    def mix(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def h(seed, r, i):
        x = mix((seed * 0x9E3779B97F4A7C15 + 0x632BE59BD9B4E019) & MASK)
        x = mix(x ^ ((r * 0xD1B54A32D192ED03) & MASK))
        return mix(x ^ ((i * 0x8CB92BA72F3D8DD7) & MASK))

    for seed, r, i in [(0, 0, 0), (0, 1, 2), (1, 63, 762111), (12345, 7, 99)]:
        assert L.orc_phase_hash(seed, r, i) == h(seed, r, i)
    v = oracle_mod.random_phase(100, 3, 0)
    assert np.allclose(np.abs(v), 1.0, atol=1e-15)


# ---------------------------------------------------------------- DoS
def test_dct3_and_jackson_closed_forms(oracle_mod):
    orc = oracle_mod
    M = 200
    g = orc.jackson(M)
    assert abs(g[0] - 1.0) < 1e-15
    rng = np.random.default_rng(1)
    mu = rng.standard_normal(M)
    a, b = 13.0, 0.3
    rho = orc.dos(mu[None, :], a, b)[0]
    k = np.arange(M)
    # FFTW REDFT01: out_k = x0 + 2 sum_{n>=1} x_n cos(pi n (k+1/2)/M)
    x = g * mu
    C_ = np.cos(np.pi * np.outer(k + 0.5, np.arange(M)) / M)
    out = x[0] + 2 * (C_[:, 1:] @ x[1:])
    E = b + a * np.cos((k + 0.5) * np.pi / M)
    want = out / (np.pi * np.sqrt(a * a - (E - b) ** 2))
    assert np.max(np.abs(rho - want)) < 1e-12
    # matlab_post/cheb_compute.m builds the same curve with an inverse DCT:
    # idct (orthonormal DCT-III) of x scaled by sqrt(2M), first term by sqrt(M)
    w = np.where(k == 0, np.sqrt(1.0 / M), np.sqrt(2.0 / M))
    idct = (C_ * w[None, :]) @ x
    scale = np.where(k == 0, 1.0, 1.0)
    assert np.allclose(out, idct * np.sqrt(2 * M) - x[0] * (np.sqrt(2) - 1) * scale, atol=1e-9)


def _golden_rows(path):
    rows = {}
    job = None
    for line in open(path):
        m = re.match(r"JOBID = (\d+)", line)
        if m:
            job = int(m.group(1))
            continue
        m = re.match(r"(\d+): (.*)", line)
        if m and job is not None:
            rows[(job, int(m.group(1)))] = np.array([float(x) for x in m.group(2).split(",")])
    return rows


@pytest.mark.parametrize("name,a", [("tblg_test1", 13.0), ("tblg_test2", 20.0)])
def test_golden_normalisation_and_oracle_structure(oracle_mod, name, a):
    """Each golden row integrates to 1 under Chebyshev-Gauss quadrature;
    the oracle reproduces the golden deck's targets, job order and the same
    normalisation (values differ: the goldens predate the current coupling
    model and ANGLE units)."""
    orc = oracle_mod
    rows = _golden_rows(golden_path(name + ".test"))
    M = 200
    x = np.cos((np.arange(M) + 0.5) * np.pi / M)
    for key, r in rows.items():
        assert len(r) == M
        assert abs(np.sum(r * np.pi * a * np.sqrt(1 - x ** 2)) / M - 1.0) < 5e-6, key
        assert np.all(r > 0)
    deck, geom, jobs, results = orc.run_deck(open(golden_path(name + ".in")).read())
    assert geom.sheet_count(0) == 115 and geom.sheet_count(1) == 115
    assert sorted({k[1] for k in rows}) == [60, 61, 175, 176]
    assert jobs[0]["target_list"] == [60, 61, 175, 176]
    if name == "tblg_test1":
        # SQ, NSHIFTS = 2 on one sheet: 4 jobs (the tblg_test2 golden was written
        # by an older LC implementation and lists 4 SQ-style jobs; the current
        # LC code makes NSHIFTS = 2 jobs, locality.cpp:1333)
        assert len(jobs) == len({k[0] for k in rows}) == 4
    else:
        assert len(jobs) == 2
    for job, rho in results:
        for r in rho:
            assert abs(np.sum(r * np.pi * a * np.sqrt(1 - x ** 2)) / M - 1.0) < 1e-12


def test_job_order_matches_golden_headers(oracle_mod):
    orc = oracle_mod
    text = open(golden_path("tblg_test1.test")).read()
    hdr = re.findall(r"2    : ([0-9.e-]+), ([0-9.e-]+), ([0-9.e-]+)", text)
    deck = orc.parse_deck(open(golden_path("tblg_test1.in")).read())
    geom = orc.Geometry(orc.geom_desc(deck))
    jobs = orc.build_jobs(deck, orc.target_list(deck, geom))
    assert [tuple(float(v) for v in h) for h in hdr] == [tuple(j["shifts"][1]) for j in jobs]
    text2 = open(golden_path("tblg_test2.test")).read()
    deck2 = orc.parse_deck(open(golden_path("tblg_test2.in")).read())
    assert "MAG_ON  = 1" in text2 and deck2.opts["magOn"] == 1 and deck2.opts["elecOn"] == 1
    geom2 = orc.Geometry(orc.geom_desc(deck2))
    jobs2 = orc.build_jobs(deck2, orc.target_list(deck2, geom2))
    # LC: shifts[last] = lc0 + x (lc1 - lc0), x = i / NSHIFTS over the default line (0,0)-(1,1)
    assert [tuple(j["shifts"][1]) for j in jobs2] == [(0.0, 0.0, 0), (0.5, 0.5, 0)]


def test_deck_errors(oracle_mod):
    orc = oracle_mod
    base = open(golden_path("tblg_test1.in")).read()
    with pytest.raises(orc.DeckError):
        orc.parse_deck(base.replace("POLY_ORDER = 200", "POLY_ORDER = 202"))
    with pytest.raises(orc.DeckError):
        orc.parse_deck(base.replace("MATERIAL = Graphene", "MATERIAL = Unobtainium"))


# ---- learned surrogate oracle --------------------------------------------------
def _mlp_oracle():
    from oracle import mlp_oracle
    return mlp_oracle


def test_mlp_oracle_gelu_closed_forms():
    O = _mlp_oracle()
    assert O.gelu(0.0) == 0.0 and O.gelu_grad(0.0) == 0.5
    assert abs(O.gelu(12.0) - 12.0) < 1e-12 and abs(O.gelu(-12.0)) < 1e-12
    z = np.linspace(-6, 6, 121)
    h = 1e-6
    fd = (O.gelu(z + h) - O.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - O.gelu_grad(z))) < 1e-8
    # the tanh form stays within 1e-3 of the exact (erf) GELU
    from scipy.special import erf
    assert np.max(np.abs(O.gelu(z) - 0.5 * z * (1 + erf(z / np.sqrt(2))))) < 1e-3


def test_mlp_oracle_bf16_rounding():
    O = _mlp_oracle()
    # exactly representable values survive; halfway cases go to the even neighbour
    assert O.bf16(np.float32(1.0)) == 1.0 and O.bf16(np.float32(-3.5)) == -3.5
    one_ulp = 2.0 ** -7                            # bf16 spacing in [1, 2)
    assert O.bf16(np.float32(1.0 + one_ulp / 2)) == 1.0                 # tie -> even (1.0)
    assert O.bf16(np.float32(1.0 + 1.5 * one_ulp)) == 1.0 + 2 * one_ulp  # tie -> even
    assert O.bf16(np.float32(1.0 + 0.6 * one_ulp)) == 1.0 + one_ulp
    x = np.random.default_rng(1).normal(size=1000).astype(np.float32)
    r = O.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)


def test_mlp_oracle_gradient_against_finite_differences():
    """Float64 loss gradient of the oracle network against central differences."""
    O = _mlp_oracle()
    rng = np.random.default_rng(2)
    dims = [6, 5, 7, 3]
    nparam = sum(dims[l + 1] * dims[l] + dims[l + 1] for l in range(len(dims) - 1))
    p = rng.normal(0, 0.5, nparam)
    X = rng.normal(size=(9, 6))
    T = rng.normal(size=(9, 3))
    _, g = O.loss_and_grad(p, dims, X, T, bf16_points=False)
    h = 1e-6
    fd = np.zeros(nparam)
    for i in range(nparam):
        e = np.zeros(nparam)
        e[i] = h
        lp, _ = O.loss_and_grad(p + e, dims, X, T, bf16_points=False)
        lm, _ = O.loss_and_grad(p - e, dims, X, T, bf16_points=False)
        fd[i] = (lp - lm) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-8 * max(1.0, np.max(np.abs(g)))


def test_mlp_features_periodic_in_shift():
    """Features of a shift are invariant under lattice translations b -> b + (1, 0), (0, 1)."""
    O = _mlp_oracle()
    rng = np.random.default_rng(4)
    cfg = np.column_stack([rng.random(20), rng.random(20), rng.normal(0, 0.003, (20, 3)), rng.uniform(0.8, 3, 20)])
    f0 = O.features(cfg)
    for d in ((1, 0), (0, 1), (-2, 3)):
        c2 = cfg.copy()
        c2[:, 0] += d[0]
        c2[:, 1] += d[1]
        assert np.max(np.abs(O.features(c2) - f0)) < 1e-5


def test_strained_hopping_closed_form(oracle_mod):
    """GrapheneSimpleStrain with HETEROSTRAIN = eps nu on the top sheet: every
    bottom-sheet bond keeps t = -2.7 eV, every top-sheet bond has the strained
    length a_cc |(I + e) R(theta) n_k| of its unit bond direction n_k and
    t = -2.7 exp(-3.37 (d / a_cc - 1)) eV, e = [[eps, 0], [0, -nu eps]]."""
    orc = oracle_mod
    text = open(os.path.join(os.path.dirname(__file__), "decks", "tblg_strain_flake.in")).read()
    od = orc.parse_deck(text)
    og = orc.Geometry(orc.geom_desc(od))
    job = orc.make_job(od, orc.build_jobs(od, orc.target_list(od, og))[0], pruned=1)
    rp, ci, v = og.assemble(job)
    pos = og.job_positions(job).reshape(-1, 3)
    resc, eps, nu, th = 10.5, 0.02, 0.16, math.radians(2.0)
    acc = 2.4768 / math.sqrt(3.0)
    E = np.array([[1 + eps, 0.0], [0.0, 1 - nu * eps]])
    Rt = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
    dirs = [np.array([math.cos(a), math.sin(a)]) for a in (math.pi / 6, 5 * math.pi / 6, -math.pi / 2)]
    top_lengths = sorted(acc * np.linalg.norm(E @ Rt @ d) for d in dirs)
    n_bottom = n_top = 0
    for r in range(len(rp) - 1):
        for q in range(rp[r], rp[r + 1]):
            c = ci[q]
            dz = pos[c, 2] - pos[r, 2]
            if c == r or dz != 0.0:
                continue
            d = float(np.linalg.norm(pos[c] - pos[r]))
            if pos[r, 2] == 0.0:
                assert abs(d - acc) < 1e-12 and abs(v[q] * resc + 2.7) < 1e-12
                n_bottom += 1
            else:
                assert min(abs(d - L) for L in top_lengths) < 1e-12, d
                assert abs(v[q] * resc + 2.7 * math.exp(-3.37 * (d / acc - 1.0))) < 1e-12
                n_top += 1
    assert n_bottom > 100 and n_top > 100
    # the three strained bond families are distinct and bracket a_cc
    assert top_lengths[0] < acc < top_lengths[-1]


def test_trace_generator_matches_libstdcxx(oracle_mod):
    """The oracle's restatement of std::default_random_engine +
    std::normal_distribution reproduces, bit for bit, the values a program built
    with this image's libstdc++ draws (tests/golden/trace_gaussians.txt, made by
    tests/golden/make_trace_gaussians.sh); minstd_rand0's 10000th output is the
    value the C++ standard requires, 1043618065."""
    orc = oracle_mod
    want = [float(line) for line in open(os.path.join(os.path.dirname(__file__), "golden", "trace_gaussians.txt"))]
    g, d = orc.MinstdRand0(), orc.NormalDistribution()
    got = [d(g) for _ in range(len(want))]
    assert got == want
    g = orc.MinstdRand0()
    for _ in range(9999):
        g()
    assert g() == 1043618065
    # generate_canonical stays in [0, 1)
    g = orc.MinstdRand0()
    u = [orc.generate_canonical(g) for _ in range(2000)]
    assert min(u) >= 0.0 and max(u) < 1.0


def test_trace_deck_normalisation(oracle_mod):
    """Trace DoS of the oracle: the Chebyshev-Gauss quadrature of rho recovers
    the sample-averaged mu_0 / N = mean |v|^2 / N of the job's Gaussian vectors."""
    orc = oracle_mod
    text = open(os.path.join(os.path.dirname(__file__), "decks", "tblg_trace.in")).read()
    deck, geom, jobs, results = orc.run_deck(text)
    o = deck.opts
    M, a = o["poly_order"], o["energy_rescale"]
    g, d = orc.MinstdRand0(), orc.NormalDistribution()
    S = orc.trace_samples(o, geom.n)
    assert S == 6 and len(results) == 4          # NSHIFTS = 2: a 2 x 2 shift grid
    for job, rho in results:
        vecs = np.array([[d(g) for _ in range(geom.n)] for _ in range(S)])
        x = np.cos((np.arange(M) + 0.5) * np.pi / M)
        quad = np.sum(rho * np.pi * a * np.sqrt(1 - x * x)) / M
        assert abs(quad - np.mean(np.sum(vecs * vecs, axis=1)) / geom.n) < 1e-10


# ---- KPM conductivity (observable type 1) ------------------------------------

def _cond_setup(orc, deck_text):
    deck = orc.parse_deck(deck_text)
    orc.setup_supercell(deck)
    geom = orc.Geometry(orc.geom_desc(deck))
    jobs = orc.build_jobs(deck, orc.target_list(deck, geom))
    return deck, geom, jobs


@pytest.mark.parametrize("periodic", [False, True])
def test_velocity_matrix_closed_form(oracle_mod, periodic):
    """dxH slots hold (x_row + image shift - x_col) t (locality.cpp:4182,4299):
    on a flake dxH = (x_i - x_j) H_ij exactly, and in both cases dxH is
    antisymmetric on H's pattern (a bond and its reverse carry opposite
    displacements)."""
    from leo2d_amd import workloads as W

    orc = oracle_mod
    text = W.tblg_cell_deck(3, 2, M=16) if periodic else W.tblg_flake_deck(3.0, 12, M=16)
    deck, geom, jobs = _cond_setup(orc, text)
    rp, ci, v, dx, dy = geom.assemble(orc.make_job(deck, jobs[0], 1), with_dx=True, with_dy=True)
    assert dx.shape == v.shape and dy.shape == v.shape
    H = _dense(rp, ci, v)
    for c, vel in enumerate((dx, dy)):
        D = _dense(rp, ci, vel)
        assert np.max(np.abs(D + D.T)) <= 1e-14 * np.max(np.abs(D))
        assert np.count_nonzero(vel) > 0.6 * len(vel)
        if not periodic:
            x = geom.index_to_pos.reshape(-1, 3)[:, c]
            assert np.max(np.abs(D - (x[:, None] - x[None, :]) * H)) == 0.0


def test_complex_velocity_is_the_k_derivative(oracle_mod):
    """H(k)_ij = sum t e^{i k.(r_j - r_i - R)}, so i dH/dk_x = sum (x_i + R_x - x_j) t e^{i phi}
    = dxH and likewise for y: the complex velocity slots (locality.cpp:4939-4940,
    5113-5114) against central differences of the oracle's own H(k) on a periodic cell."""
    from leo2d_amd import workloads as W

    orc = oracle_mod
    deck, geom, jobs = _cond_setup(orc, W.tblg_cell_deck(3, 2, M=16))
    k0 = np.array([0.0317, -0.0211])

    def hk(k, vel=False):
        j = orc.make_job(deck, jobs[0], 0)
        j.complex_matrix = 1
        j.k[0], j.k[1] = k
        out = geom.assemble(j, with_dx=vel, with_dy=vel)
        return out if vel else out[2]

    rp, ci, v, dx, dy = hk(k0, True)
    assert np.iscomplexobj(dx) and np.max(np.abs(dx.imag)) > 0
    h = 1e-5
    for e, vel in ((np.array([h, 0.0]), dx), (np.array([0.0, h]), dy)):
        deriv = (hk(k0 + e) - hk(k0 - e)) / (2 * h)
        assert np.max(np.abs(1j * deriv - vel)) <= 1e-7 * np.max(np.abs(vel))


def test_cond_against_dense_eigendecomposition(oracle_mod):
    """C[n][m] = (dxH T_n(H) alpha_0) . (T_m(H) e_t), alpha_0 = -dxH e_t, with the
    T_n built from the eigendecomposition of the dense H instead of the
    recursion."""
    from leo2d_amd import workloads as W

    orc = oracle_mod
    deck, geom, jobs = _cond_setup(orc, W.tblg_flake_deck(3.0, 12, M=16))
    rp, ci, v, dx = geom.assemble(orc.make_job(deck, jobs[0], 1), with_dx=True)
    H, D = _dense(rp, ci, v), _dense(rp, ci, dx)
    lam, U = np.linalg.eigh(H)
    assert np.max(np.abs(lam)) < 1.0
    M = 12
    targets = jobs[0]["target_list"][:3]
    Cm = orc.cond(rp, ci, v, dx, M, targets)
    Tn = [(U * np.cos(k * np.arccos(lam))) @ U.T for k in range(M)]
    for r, t in enumerate(targets):
        e = np.zeros(geom.n)
        e[t] = 1.0
        a0 = -D @ e
        exact = np.array([[(D @ Tn[j] @ a0) @ (Tn[m] @ e) for m in range(M)] for j in range(M)])
        assert np.max(np.abs(Cm[r] - exact)) <= 1e-13 * np.max(np.abs(exact))
        # C[0][0] = alpha_0 . dxH^T... = |dxH e_t|^2 up to sign: (dxH a0) . e = -(dxH dxH e)_t = |dxH e|^2
        assert abs(Cm[r, 0, 0] - np.sum((D @ e) ** 2)) <= 1e-14 * max(1.0, abs(Cm[r, 0, 0]))


def test_cond_transform_against_scipy_dctn(oracle_mod):
    """matrixResponseTransform is FFTW's 2D REDFT01 of 2 g_i g_j C_ij with the
    Jackson g_i divided by sqrt(2); scipy's unnormalised DCT-III has the same
    definition (y_k = x_0 + 2 sum_{n>=1} x_n cos(pi n (2k+1) / 2N))."""
    from scipy.fft import dctn

    orc = oracle_mod
    rng = np.random.default_rng(3)
    for M in (8, 33):
        Cm = rng.standard_normal((M, M))
        out = orc.cond_transform(Cm)
        i = np.arange(M)
        g = ((M - i) * np.cos(np.pi * i / M) + np.sin(np.pi * i / M) / np.tan(np.pi / M)) / (np.sqrt(2.0) * M)
        want = dctn(2.0 * np.outer(g, g) * Cm, type=3)
        assert np.max(np.abs(out - want)) <= 1e-12 * np.max(np.abs(want)), M


# ---------------------------------------------------------------- k-point paths
DECKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "decks")


def _k_deck(orc, name, k_type, nk, *edits):
    text = open(os.path.join(DECKS, name)).read()
    text = re.sub(r"K_TYPE = \d+", "K_TYPE = %d" % k_type, text)
    text = re.sub(r"K_GRID = \d+ \d+", "K_GRID = %d 1" % nk, text)
    for a, b in edits:
        text = text.replace(a, b)
    od = orc.parse_deck(text)
    orc.setup_supercell(od)
    return od


def _recip(a):
    """Rows b_i with b_i . a_j = 2 pi delta_ij."""
    return 2 * math.pi * np.linalg.inv(np.array(a, dtype=float)).T


def _rot(v, deg):
    t = math.radians(deg)
    return np.array([math.cos(t) * v[0] - math.sin(t) * v[1], math.sin(t) * v[0] + math.cos(t) * v[1]])


def _corner(p, q):
    """Circumcentre of the triangle 0, p, q: the x with 2 x.p = |p|^2 and 2 x.q = |q|^2."""
    return np.linalg.solve(2.0 * np.array([p, q]), [p @ p, q @ q])


def _walk(corners, span, nk):
    """Points at c = span i / nk along corners[0] -> corners[1] -> ..., one leg per unit of c."""
    out = []
    for i in range(nk):
        c = span * i / nk
        s = min(int(math.floor(c)), len(corners) - 2)
        out.append(corners[s] + (c - s) * (corners[s + 1] - corners[s]))
    return np.array(out)


def _is_zone_corner(K, b):
    """K is a vertex of the Wigner-Seitz cell of the lattice spanned by b: three
    lattice points (the origin one of them) at the same, smallest distance."""
    d = sorted(np.hypot(*(K - (n1 * b[0] + n2 * b[1])))
               for n1 in range(-3, 4) for n2 in range(-3, 4))
    return abs(d[0] - np.hypot(*K)) < 1e-12 * d[0] and abs(d[2] - d[0]) < 1e-12 * d[0] and d[3] > d[0] * 1.01


def test_reciprocal_against_matrix_inverse(oracle_mod):
    """reciprocal(a) and the sheet lattice's reciprocal (turned, uniformly strained)
    satisfy b_i . a_j = 2 pi delta_ij."""
    od = _k_deck(oracle_mod, "tblg_sc_7_6_k.in", 0, 2)
    assert np.allclose(oracle_mod.reciprocal(od.opts["supercell"]), _recip(od.opts["supercell"]), rtol=0, atol=1e-14)
    od = _k_deck(oracle_mod, "mlg_strain_k2.in", 2, 9, ("0.00 0.05 0.05 0.00", "0.01 0.05 -0.02 0.03"))
    u = np.array(od.opts["u_ij"])
    assert u[0][1] != u[1][0]
    a = np.array(oracle_mod.LATTICE) @ (np.eye(2) + u).T          # a_i -> (1 + u) a_i
    assert np.allclose(oracle_mod.sheet_reciprocal(od, 0), _recip(a), rtol=0, atol=1e-14)
    od = _k_deck(oracle_mod, "ttlg_grid_k5.in", 5, 4)
    a = np.array([_rot(v, 2.1) for v in oracle_mod.LATTICE])
    assert np.allclose(oracle_mod.sheet_reciprocal(od, 2), _recip(a), rtol=0, atol=1e-14)


def _path(orc, od):
    return np.array(orc.k_path(od))


@pytest.mark.parametrize("nk", [6, 7])
def test_k_type1_supercell_path(oracle_mod, nk):
    """K -> Gamma -> M -> K of the (7, 6) cell's zone: K is the zone corner next to
    M = b2 / 2 on the clockwise side, M the midpoint of the edge facing b2."""
    od = _k_deck(oracle_mod, "tblg_sc_7_6_k.in", 1, nk)
    b = _recip(od.opts["supercell"])
    K = _corner(b[1], _rot(b[1], -60))
    assert _is_zone_corner(K, b)
    want = _walk([K, np.zeros(2), b[1] / 2, K], 3.0, nk)
    got = _path(oracle_mod, od)
    assert got.shape == (nk, 2)
    assert np.allclose(got, want, rtol=0, atol=1e-12 * np.hypot(*b[1]))


def test_k_type2_monolayer_path(oracle_mod):
    """Gamma -> M -> K -> Gamma of graphene's own zone: M = b1 / 2 and K the corner
    at 4 pi / (3 a) (textbook), i.e. K = (4 pi / 3 a, 0) for the unturned sheet."""
    od = _k_deck(oracle_mod, "mlg_strain_k2.in", 2, 9, ("UNIFORM_STRAIN = 1", "UNIFORM_STRAIN = 0"))
    b = _recip(oracle_mod.LATTICE)
    K = _corner(b[0], _rot(b[0], 60))
    assert _is_zone_corner(K, b)
    assert np.allclose(K, [4 * math.pi / (3 * oracle_mod.GR_A), 0.0], rtol=0, atol=1e-14)
    got = _path(oracle_mod, od)
    assert np.allclose(got, _walk([np.zeros(2), b[0] / 2, K, np.zeros(2)], 3.0, 9), rtol=0, atol=1e-13)


def test_k_type3_supercell_path(oracle_mod):
    """Gamma -> M -> K -> K' -> Gamma: K' is the next zone corner clockwise."""
    od = _k_deck(oracle_mod, "tblg_sc_7_6_k.in", 3, 8)
    b = _recip(od.opts["supercell"])
    K1 = _corner(b[1], _rot(b[1], -60))
    K2 = _corner(_rot(b[1], -60), _rot(b[1], -120))
    assert _is_zone_corner(K1, b) and _is_zone_corner(K2, b)
    want = _walk([np.zeros(2), b[1] / 2, K1, K2, np.zeros(2)], 4.0, 8)
    assert np.allclose(_path(oracle_mod, od), want, rtol=0, atol=1e-12 * np.hypot(*b[1]))


@pytest.mark.parametrize("mn", ["7 6", "6 7"])
def test_k_type4_cell_path(oracle_mod, mn):
    """Gamma -> K -> M -> Gamma of the (M, N) cell built on the unturned sheet
    lattice, nk points per line from each line's start."""
    nk = 5
    od = _k_deck(oracle_mod, "tblg_sc_7_6_k.in", 4, nk, ("SUPERCELL_M_N = 7 6", "SUPERCELL_M_N = " + mn))
    M, N = (int(x) for x in mn.split())
    A = np.array([[N, M], [-M, M + N]] if M > N else [[M, N], [-N, M + N]], dtype=float)
    b = _recip(A @ np.array(oracle_mod.LATTICE))
    K = _corner(b[1], _rot(b[1], -60))
    Mp = (b[0] + b[1]) / 2
    assert _is_zone_corner(K, b)
    corners = [np.zeros(2), K, Mp, np.zeros(2)]
    want = np.array([corners[i // nk] + (i % nk) / nk * (corners[i // nk + 1] - corners[i // nk])
                     for i in range(3 * nk)])
    assert np.allclose(_path(oracle_mod, od), want, rtol=0, atol=1e-12 * np.hypot(*b[1]))


def test_k_type5_trilayer_dirac_points(oracle_mod):
    """K of sheet 1 -> K of sheet 2 -> K of sheet 3: each K is a corner of its turned
    sheet's zone at 4 pi / (3 a); 2 nk points on the two lines, then K3."""
    nk = 4
    od = _k_deck(oracle_mod, "ttlg_grid_k5.in", 5, nk)
    Ks = []
    for s in range(3):
        b = _recip([_rot(v, math.degrees(od.angles[s])) for v in oracle_mod.LATTICE])
        K = _corner(b[0], b[0] + b[1])
        assert _is_zone_corner(K, b)
        assert abs(np.hypot(*K) - 4 * math.pi / (3 * oracle_mod.GR_A)) < 1e-14
        Ks.append(K)
    want = [Ks[0] + i / nk * (Ks[1] - Ks[0]) for i in range(nk)]
    want += [Ks[1] + i / nk * (Ks[2] - Ks[1]) for i in range(nk)] + [Ks[2]]
    assert np.allclose(_path(oracle_mod, od), np.array(want), rtol=0, atol=1e-13)


def test_kpts_text_format(oracle_mod):
    """saveKpts: std::fixed, precision 10, a space before a value >= 0 (so before
    -0.0 too), two tabs, ' \\n'; jobs without a Bloch vector write nothing."""
    jobs = [{"k_vec": [0.5, -0.25]}, {"k_vec": [-0.0, 1e-17]}, {"k_vec": [-1e-12, 3.0]}, {"jobID": 4}]
    assert oracle_mod.kpts_text(jobs) == (" 0.5000000000\t\t-0.2500000000 \n"
                                          " -0.0000000000\t\t 0.0000000000 \n"
                                          "-0.0000000000\t\t 3.0000000000 \n")


# ---------------------------------------------------------------- relaxation field
def _relax_job(orc, coeffs, max_k, b):
    j = orc.Job()
    orc.set_relaxation(j, max_k, np.asarray(coeffs, dtype=float).reshape(3, -1), b)
    return j


def _cell_b(orc):
    from leo2d_amd import workloads as W
    deck = orc.parse_deck(W.tblg_cell_deck(7, 6, M=16))
    orc.setup_supercell(deck)
    return np.array(deck.opts["supercell"]), np.array(orc.reciprocal(deck.opts["supercell"]))


def test_relaxation_closed_forms(oracle_mod):
    """StrainCalc::fourierStrainDisp_sc on hand-set coefficients:
    * the constant term moves every orbital by c_z0 out of plane only;
    * at the origin every sine vanishes and u_z = c_z0 + 6 sum_{idx >= 1} c_z[idx]
      (three turns, factor 2 for e^{ikr} + c.c.);
    * a single (1, 0) term with c along the supercell basis's x (that is, along
      b1) is longitudinal: u = 2 c sum_r (k_r / |k_r|) sin(k_r . r) over k_r = R(r 60 deg) b1;
    * the second sheet moves the opposite way."""
    orc = oracle_mod
    A, b = _cell_b(orc)
    j = _relax_job(orc, [0.0, 0.0, 0.37], 0, b)
    assert np.array_equal(orc.relax_disp(j, 3.1, -7.2, 0), [0.0, 0.0, 0.37])
    assert np.array_equal(orc.relax_disp(j, 3.1, -7.2, 1), [0.0, 0.0, -0.37])
    cz = [0.11, -0.05, 0.02, 0.013]
    j = _relax_job(orc, [[0.3, 0.1, -0.2, 0.05], [0.0, 0.2, 0.1, -0.1], cz], 2, b)
    d = orc.relax_disp(j, 0.0, 0.0, 0)
    assert np.allclose(d, [0.0, 0.0, cz[0] + 6 * sum(cz[1:])], rtol=0, atol=1e-15)
    c = 0.043
    j = _relax_job(orc, [[0.0, c], [0.0, 0.0], [0.0, 0.0]], 1, b)
    rng = np.random.default_rng(5)
    for x, y in rng.uniform(-40, 40, (20, 2)):
        want = np.zeros(2)
        for r in range(3):
            k = _rot(b[0], 60 * r)
            want += 2 * c * k / np.hypot(*k) * math.sin(k @ [x, y])
        got = orc.relax_disp(j, x, y, 0)
        assert np.allclose(got[:2], want, rtol=0, atol=1e-14) and got[2] == 0.0
        assert np.allclose(orc.relax_disp(j, x, y, 1), -got, rtol=0, atol=0)


def test_relaxation_is_periodic_and_threefold(oracle_mod):
    """With every coefficient set, the field repeats with the supercell and turns
    with the lattice: u(r + A_i) = u(r), u_z(R120 r) = u_z(r) and
    u_xy(R120 r) = R120 u_xy(r)."""
    orc = oracle_mod
    A, b = _cell_b(orc)
    mk, co = orc.read_relaxation(os.path.join(DECKS, "tblg_relax_coeffs.txt"))
    co = co + np.random.default_rng(1).normal(0, 0.01, co.shape)     # fill the zero rows too
    j = _relax_job(orc, co, mk, b)
    rng = np.random.default_rng(2)
    for x, y in rng.uniform(-30, 30, (25, 2)):
        u = orc.relax_disp(j, x, y, 0)
        for a in A:
            assert np.allclose(orc.relax_disp(j, x + a[0], y + a[1], 0), u, rtol=0, atol=1e-12)
        p = _rot(np.array([x, y]), 120)
        v = orc.relax_disp(j, p[0], p[1], 0)
        assert abs(v[2] - u[2]) < 1e-12
        assert np.allclose(v[:2], _rot(u[:2], 120), rtol=0, atol=1e-12)


def test_local_twist_of_an_unrelaxed_sheet(oracle_mod):
    """getLocalTheta on unrelaxed positions: every nearest-neighbour bond of a sheet
    turned by theta gives theta back, modulo the 120-degree bond star; an orbital
    with no neighbour at index offsets -3, -1, 1, 3 takes the sheet angle."""
    from leo2d_amd import workloads as W

    orc = oracle_mod
    for text in (W.tblg_cell_deck(7, 6, M=16), W.tblg_flake_deck(3.0, 12, M=16)):
        deck = orc.parse_deck(text)
        orc.setup_supercell(deck)
        geom = orc.Geometry(orc.geom_desc(deck))
        th = geom.local_theta(geom.index_to_pos)
        sheet = geom.index_to_grid.reshape(-1, 4)[:, 3]
        ang = np.array(deck.angles)[sheet]
        third = 2 * math.pi / 3
        dev = np.remainder(th - ang + third / 2, third) - third / 2
        assert np.max(np.abs(dev)) < 1e-12
        assert np.mean(th != ang) > 0.5                 # most orbitals found a bond


def test_direct_plane_wave_field(oracle_mod):
    """STRAIN_TYPE = DIRECT (StrainCalc::supercellDisp): the same Fourier sum over
    the reference's built-in max_k = 2 coefficients, scaled per sheet.  At the
    origin u_z = c_z0 + 6 (c_z1 + c_z2 + c_z3) on sheet 1; sheet 2 moves in plane
    by -2 times sheet 1 and not at all out of plane; sheet 3 mirrors sheet 1 out of
    plane; the field repeats with the supercell."""
    orc = oracle_mod
    A, b = _cell_b(orc)
    j = orc.Job()
    j.strain_type = 6
    for a in range(2):
        for d in range(2):
            j.relax_b[a][d] = b[a][d]
    cz = [-0.061901691526695, -0.0220634404118861, -0.004100201789620, -0.002676762110702]
    d0 = orc.relax_disp(j, 0.0, 0.0, 0)
    assert np.allclose(d0, [0.0, 0.0, cz[0] + 6 * sum(cz[1:])], rtol=0, atol=1e-15)
    rng = np.random.default_rng(3)
    for x, y in rng.uniform(-30, 30, (20, 2)):
        u0, u1, u2 = (orc.relax_disp(j, x, y, s) for s in range(3))
        assert np.hypot(*u0[:2]) > 0
        assert np.array_equal(u1[:2], -2 * u0[:2]) and u1[2] == 0.0
        assert np.array_equal(u2[:2], u0[:2]) and u2[2] == -u0[2]
        for a in A:
            assert np.allclose(orc.relax_disp(j, x + a[0], y + a[1], 0), u0, rtol=0, atol=1e-12)


# ------------------------------------------------- field terms, supercell, slots
def _deck_with(text, **keys):
    """The deck text with KEY = value lines replaced (or appended before POLY_ORDER)."""
    for k, v in keys.items():
        if re.search(r"^%s\b.*$" % k, text, flags=re.M):
            text = re.sub(r"^%s\b.*$" % k, "%s = %s" % (k, v), text, flags=re.M)
        else:
            text = text.replace("POLY_ORDER", "%s = %s\nPOLY_ORDER" % (k, v), 1)
    return text


def _entry(rp, ci, v, i, j):
    q = [p for p in range(rp[i], rp[i + 1]) if ci[p] == j]
    assert len(q) == 1, (i, j)
    return v[q[0]]


def _hexagons(pos, sheet_of, sheet, centres_near, acc=1.42):
    """Hexagons of one graphene sheet: the six orbitals within acc + 0.01 A of a
    ring centre, ordered counterclockwise.  Centres are found geometrically:
    an orbital's three bonds point away from three ring centres at distance acc."""
    sel = np.where(sheet_of == sheet)[0]
    p2 = pos[sel, :2]
    out = []
    for c0 in centres_near:
        # the orbital nearest to c0, then the centre opposite one of its bonds
        a = sel[np.argmin(np.sum((p2 - c0) ** 2, axis=1))]
        d = np.sqrt(np.sum((p2 - pos[a, :2]) ** 2, axis=1))
        nb = sel[(d > 1.0) & (d < 1.6)]
        c = pos[a, :2] - (pos[nb[0], :2] - pos[a, :2])
        ring = sel[np.abs(np.sqrt(np.sum((p2 - c) ** 2, axis=1)) - acc) < 0.01]
        assert len(ring) == 6, len(ring)
        ang = np.arctan2(pos[ring, 1] - c[1], pos[ring, 0] - c[0])
        out.append((c, ring[np.argsort(ang)]))
    return out


def test_peierls_plaquette_flux(oracle_mod):
    """Uniform B (locality.cpp:7157-7161, peierlsPhase): the product of H around
    any graphene hexagon, taken counterclockwise, has phase -3.14e-5 B A
    (A = the hexagon's area from the orbital positions, shoelace formula) and
    modulus |t_NN / a|^6; gauge invariance makes that the same for every
    hexagon of the flake, on both sheets.  A wrong prefactor or sign, or a
    phase that is not the trapezoid (x2 - x1)(y1 + y2)/2 (e.g. y1 alone), fails
    one of these.  With B = 0 the same products are real and positive."""
    orc = oracle_mod
    B, a = 500.0, 13.0
    base = open(golden_path("tblg_test1.in")).read()
    for Bf, on in ((B, 1), (0.0, 0)):
        text = _deck_with(base, USE_B_FIELD=on, B_FIELD=Bf)
        deck, geom, jobs, _ = orc.run_deck(text)
        job = orc.make_job(deck, jobs[0])
        assert job.complex_matrix == on
        rp, ci, v = geom.assemble(job)
        pos = orc.Geometry.job_positions(geom, job).reshape(-1, 3)
        sheet_of = geom.index_to_grid.reshape(-1, 4)[:, 3]
        for s in (0, 1):
            for c, ring in _hexagons(pos, sheet_of, s, [np.array([0.0, 0.0]), np.array([4.1, -2.3]),
                                                         np.array([-5.0, 3.7])]):
                prod = 1.0 + 0j
                for k in range(6):
                    prod *= _entry(rp, ci, v, ring[k], ring[(k + 1) % 6])
                x, y = pos[ring, 0], pos[ring, 1]
                area = 0.5 * np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)
                assert area > 5.0
                want = -3.14e-5 * Bf * area
                assert abs(np.angle(prod) - want) <= 1e-12, (s, c, np.angle(prod), want)
                assert abs(abs(prod) - (2.8922 / a) ** 6) <= 1e-12 * (2.8922 / a) ** 6


def test_electric_field_onsite(oracle_mod):
    """USE_E_FIELD (locality.cpp:4166-4176, onSiteE = z E at :7163-7166): every
    diagonal entry is (0.3504 + ENERGY_SHIFT + z E) / ENERGY_RESCALE, z the
    orbital's height (0 on sheet 1, HEIGHT of sheet 2); off-diagonal entries
    do not move."""
    orc = oracle_mod
    base = open(golden_path("tblg_test1.in")).read()
    E, shift, a = 0.7, 0.25, 13.0
    out = {}
    for on in (0, 1):
        text = _deck_with(base, USE_E_FIELD=on, E_FIELD=E if on else 0, ENERGY_SHIFT=shift, ENERGY_RESCALE=a)
        deck, geom, jobs, _ = orc.run_deck(text)
        job = orc.make_job(deck, jobs[0])
        out[on] = geom.assemble(job)
        z = geom.index_to_pos.reshape(-1, 3)[:, 2]
    rp, ci, v1 = out[1]
    _, _, v0 = out[0]
    assert np.array_equal(out[0][0], rp) and np.array_equal(out[0][1], ci)
    heights = sorted(set(z.tolist()))
    assert heights == [0.0, 3.4]
    for i in range(len(rp) - 1):
        for q in range(rp[i], rp[i + 1]):
            if ci[q] == i:
                assert v1[q] == (0.3504 + shift + z[i] * E) / a
                assert v0[q] == (0.3504 + shift) / a
            else:
                assert v1[q] == v0[q]


@pytest.mark.parametrize("mn,deg", [((7, 6), 5.0858), ((6, 7), -5.0858), ((32, 31), 1.0501)])
def test_supercell_angle_and_commensurability(oracle_mod, mn, deg):
    """setupSupercell type 1 (locality.cpp:131-289).  The twist is measured from
    the orbital positions (the angle between the two sheets' lattice vector a1)
    and compared with the published commensurate angles of the (i+1, i) cells
    (5.0858 deg for i = 6, 1.0501 deg for i = 31; negative when M < N).  The
    supercell is commensurate: N a1 + M a2 of sheet 1 equals M a1 + N a2 of
    sheet 2, and -M a1 + (M+N) a2 equals -N a1 + (M+N) a2, each sheet's a1, a2
    read off neighbouring grid cells of its positions."""
    orc = oracle_mod
    M, N = mn
    from leo2d_amd import workloads as W
    text = W.tblg_cell_deck(M, N, 1)
    od = orc.parse_deck(text)
    orc.setup_supercell(od)
    og = orc.Geometry(orc.geom_desc(od))
    g = og.index_to_grid.reshape(-1, 4)
    pos = og.index_to_pos.reshape(-1, 3)
    assert og.n == 4 * (N * N + N * M + M * M)
    where = {tuple(r): k for k, r in enumerate(g.tolist())}
    vec = {}
    for s in (0, 1):
        for d, key in (((1, 0), "a1"), ((0, 1), "a2")):
            for (i, j, l, ss), k in where.items():
                if ss == s and l == 0 and (i + d[0], j + d[1], 0, s) in where:
                    vec[s, key] = pos[where[(i + d[0], j + d[1], 0, s)], :2] - pos[k, :2]
                    break
    ang = math.degrees(math.atan2(vec[1, "a1"][1], vec[1, "a1"][0]) - math.atan2(vec[0, "a1"][1], vec[0, "a1"][0]))
    assert abs(ang - deg) < 1e-4, ang
    A1 = N * vec[0, "a1"] + M * vec[0, "a2"]
    A2 = -M * vec[0, "a1"] + (M + N) * vec[0, "a2"]
    B1 = M * vec[1, "a1"] + N * vec[1, "a2"]
    B2 = -N * vec[1, "a1"] + (M + N) * vec[1, "a2"]
    assert np.max(np.abs(A1 - B1)) < 1e-9 and np.max(np.abs(A2 - B2)) < 1e-9
    sc = np.array(od.opts["supercell"])
    assert np.max(np.abs(sc[0] - A1)) < 1e-9 and np.max(np.abs(sc[1] - A2)) < 1e-9


def _brute_csr(orc, od, og, job):
    """generateRealH / generateCpxH slot rule restated pair by pair from the
    pair lists (locality.cpp:4120-4192 intra, 4220-4309 inter; complex
    4840-4975, 5060-5140): per row, the intralayer pairs in list order, then
    the interlayer pairs; a run of consecutive pairs with the same (row,
    column) is one slot (periodic images of one column), a slot is stored even
    when its value is zero, and each pair adds t / ENERGY_RESCALE, the
    diagonal pairs (every image for a real H, the home image for a complex
    one) adding (t + ENERGY_SHIFT + z E) / ENERGY_RESCALE; complex entries are
    polar(t, 3.14e-5 B (x2 - x1)(y1 + y2)/2 + k.(r2 - r1)), r1 the row orbital
    moved by the pair's supercell image."""
    L = orc.lib()
    n = og.n
    pos = og.job_positions(job).reshape(-1, 3)
    g = og.index_to_grid.reshape(-1, 4)
    sc = np.array(od.opts.get("supercell", [[0.0, 0.0], [0.0, 0.0]]))
    angles = od.angles
    mats = [sd.mat for sd in od.sdata]
    cpx = bool(job.complex_matrix)
    a, shift = job.energy_rescale, job.energy_shift
    blocks = []
    for which in (0, 1):
        ij, ims, t = og.pair_arrays(which)
        blocks.append((ij, ims, t))
    rp = [0]
    cols, vals = [], []
    cur = [0, 0]
    for i in range(n):
        for which in (0, 1):
            ij, ims, t = blocks[which]
            q = cur[which]
            acc, open_col = None, None
            while q < len(ij) and ij[q, 0] == i:
                j = int(ij[q, 1])
                si, sj = int(ims[q, 0]), int(ims[q, 1])
                x1 = pos[i, 0] + si * sc[0, 0] + sj * sc[1, 0]
                y1 = pos[i, 1] + si * sc[0, 1] + sj * sc[1, 1]
                z1 = pos[i, 2]
                x2, y2, z2 = pos[j]
                if which == 0:
                    raw = float(t[q])
                else:
                    raw = L.orc_interlayer(mats[g[i, 3]], int(g[i, 2]), int(g[j, 2]), x2 - x1, y2 - y1, z2 - z1,
                                           angles[g[i, 3]], angles[g[j, 3]])
                diag = j == i and (not cpx or (si == 0 and sj == 0))
                if diag:
                    tv = (raw + shift + z1 * job.E) / a if job.elec_on == 1 else (raw + shift) / a
                else:
                    tv = raw / a
                if cpx:
                    ph = 3.14e-5 * job.B * (x2 - x1) * 0.5 * (y1 + y2) + job.k[0] * (x2 - x1) + job.k[1] * (y2 - y1)
                    tv = complex(tv * math.cos(ph), tv * math.sin(ph))
                if acc is None:
                    acc = 0j if cpx else 0.0
                if tv != 0:
                    acc += tv
                nxt_same = q + 1 < len(ij) and ij[q + 1, 0] == i and ij[q + 1, 1] == j
                if not nxt_same:
                    cols.append(j)
                    vals.append(acc)
                    acc = None
                q += 1
            cur[which] = q
        rp.append(len(cols))
    return np.array(rp), np.array(cols, dtype=np.int32), np.array(vals)


@pytest.mark.parametrize("case", ["flake", "flake_fields", "grid_real", "grid_k"])
def test_slot_rule_against_pairwise_restatement(oracle_mod, case):
    """The CSR pattern and values of the oracle's assembly equal a pair-by-pair
    restatement of the reference's slot rule built from the pinned pair lists
    and couplings, on a flake (no images), a flake in B and E fields, and the
    3 x 3 grid cell where one row meets the same column through several
    supercell images (merged into one slot), real and at a k-point.  Dropping zero-valued slots, merging
    non-consecutive repeats, putting the interlayer block first or moving the
    diagonal shift each change the pattern or the values."""
    orc = oracle_mod
    if case.startswith("flake"):
        text = open(golden_path("tblg_test1.in")).read()
        if case == "flake_fields":
            text = _deck_with(text, USE_B_FIELD=1, B_FIELD=300, USE_E_FIELD=1, E_FIELD=0.4, ENERGY_SHIFT=0.3)
    else:
        text = open(os.path.join(os.path.dirname(golden_path("x")), "..", "decks", "tblg_grid_3x3.in")).read()
    od = orc.parse_deck(text)
    orc.setup_supercell(od)
    og = orc.Geometry(orc.geom_desc(od))
    jobs = orc.build_jobs(od, orc.target_list(od, og))
    job = orc.make_job(od, jobs[-1])
    if case == "grid_real":
        job.complex_matrix = 0
        job.k[0], job.k[1] = 0.0, 0.0
    if case == "grid_k":
        assert job.complex_matrix == 1 and (job.k[0] != 0 or job.k[1] != 0)
    rp, ci, v = og.assemble(job)
    brp, bci, bv = _brute_csr(orc, od, og, job)
    assert np.array_equal(rp, brp)
    assert np.array_equal(ci, bci)
    assert np.max(np.abs(v - bv)) <= 1e-15 * max(1.0, float(np.max(np.abs(bv))))
    if case == "grid_real":
        # the grid cell does exercise the rule: runs of one column through several
        # images (fewer slots than pairs)
        npairs = sum(len(og.pair_arrays(w)[0]) for w in (0, 1))
        assert npairs > len(ci) + 1000
    if case == "flake":
        # ... and the flake keeps zero-coupling home-image slots
        assert np.sum(v == 0) > 100


@pytest.mark.parametrize("deckname,vec", [("tblg_test1.in", "real"), ("tblg_test2.in", "phase")])
def test_kpm_loop_matches_reference_loop(oracle_mod, ref_lib, deckname, vec):
    """The oracle's faithful KPM (orc_kpm, reference scheme with vector starts)
    equals, bit for bit, the reference's own loop around its own
    SpMatrix::vectorMultiply (oracle/_ref ref_kpm_time, locality.cpp:5990-6041,
    6138-6346), run on three threads: real H with real vectors, complex H(B)
    with random-phase vectors.  The same library call is bench.py's CPU
    baseline."""
    orc = oracle_mod
    L = ref_lib
    L.ref_kpm_time.restype = C.c_double
    L.ref_kpm_time.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                               C.c_int, C.c_int, C.c_int, C.c_void_p]
    deck, geom, jobs, _ = orc.run_deck(open(golden_path(deckname)).read())
    rp, ci, v = geom.assemble(orc.make_job(deck, jobs[0]))
    n, M, R = geom.n, 120, 5
    cpx = np.iscomplexobj(v)
    vecs = orc.random_phase(n, R, 3) if vec == "phase" else np.random.default_rng(1).standard_normal((R, n))
    mu = np.zeros((R, M))
    rp32 = np.ascontiguousarray(rp, dtype=np.int32)
    ci32 = np.ascontiguousarray(ci, dtype=np.int32)
    vv = np.ascontiguousarray(v.view(np.float64) if cpx else v)
    vf = np.ascontiguousarray(vecs).view(np.float64) if np.iscomplexobj(vecs) else np.ascontiguousarray(vecs)
    assert np.iscomplexobj(vecs) == cpx
    L.ref_kpm_time(n, len(ci), _p(rp32), _p(ci32), _p(vv), int(cpx), _p(vf), R, M, 3, _p(mu))
    omu = orc.kpm(rp, ci, v, M, vecs=vecs, nthreads=2)
    assert np.array_equal(mu, omu)
