"""KPM conductivity (OBSERVABLE_TYPE = 1) on the device against the oracle.

The oracle restates computeCondKPM (locality.cpp:6348-6727) and
matrixResponseTransform (param_tools.cpp:838-905) and is pinned by a dense
eigendecomposition and scipy's DCT-III (tests/test_oracle_pins.py).  Bars:
the velocity matrix's pattern bit-exact and its values within 1e-12 of the
largest; moments within 1e-12 of the largest (FP64; the matrix-core GEMM sums rows in
another order than the reference's dot loop); the transform within 1e-12 of
its largest value.
"""
import os

import numpy as np
import pytest

import leo2d_amd as L
from leo2d_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = L.Device(0)
    yield d
    d.close()


def _oracle_setup(orc, text, pruned):
    deck = orc.parse_deck(text)
    orc.setup_supercell(deck)
    geom = orc.Geometry(orc.geom_desc(deck))
    jobs = orc.build_jobs(deck, orc.target_list(deck, geom))
    return geom.assemble(orc.make_job(deck, jobs[0], pruned), with_dx=True, with_dy=True), jobs[0]["target_list"]


def _device_setup(dev, text, pattern):
    geom = L.Geometry(L.Deck(text))
    dev.set_geometry(geom)
    dev.pairs_build()
    job = geom.jobs()[0]
    dev.assemble_velocity(job, pattern)
    return geom, job


DECKS = [("flake", W.tblg_flake_deck(3.0, 14, M=16)), ("cell", W.tblg_cell_deck(3, 2, M=16))]


@pytest.mark.parametrize("name,text", DECKS)
@pytest.mark.parametrize("pattern", [L.PATTERN_PRUNED, L.PATTERN_REFERENCE])
def test_velocity_matrix_matches_oracle(dev, oracle_mod, name, text, pattern):
    print("case", name, pattern, flush=True)
    (orp, oci, ov, odx, ody), _ = _oracle_setup(oracle_mod, text, 1 if pattern == L.PATTERN_PRUNED else 0)
    _device_setup(dev, text, pattern)
    rp, ci, v = dev.csr_copy()
    dx, dy = dev.velocity_copy(with_dy=True)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    assert np.max(np.abs(v - ov)) <= 1e-12 * np.max(np.abs(ov))
    assert np.max(np.abs(dx - odx)) <= 1e-12 * np.max(np.abs(odx))
    assert np.max(np.abs(dy - ody)) <= 1e-12 * np.max(np.abs(ody))
    assert np.array_equal(dev.velocity_copy(), dx)


@pytest.mark.parametrize("pattern", [L.PATTERN_PRUNED, L.PATTERN_REFERENCE])
def test_complex_velocity_matrices_match_oracle(dev, oracle_mod, pattern):
    """For a complex H(k) the velocity slots carry the phased coupling:
    dxH = sum (x1 - x2) t e^{i phi}, likewise dyH (locality.cpp:4939-4940, 5113-5114)."""
    print("case complex velocity", pattern, flush=True)
    text = open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "decks", "tblg_sc_7_6_k.in")).read()
    deck = oracle_mod.parse_deck(text)
    oracle_mod.setup_supercell(deck)
    og = oracle_mod.Geometry(oracle_mod.geom_desc(deck))
    ojobs = oracle_mod.build_jobs(deck, oracle_mod.target_list(deck, og))
    q = 3                                           # k = (b1 + b2) / 2: every phase non-trivial
    orp, oci, ov, odx, ody = og.assemble(oracle_mod.make_job(deck, ojobs[q], 1 if pattern == L.PATTERN_PRUNED else 0),
                                         with_dx=True, with_dy=True)
    geom = L.Geometry(L.Deck(text))
    dev.set_geometry(geom)
    dev.pairs_build()
    job = geom.jobs()[q]
    assert job.complex_matrix == 1 and (job.k[0], job.k[1]) == tuple(ojobs[q]["k_vec"])
    dev.assemble_velocity(job, pattern)
    rp, ci, v = dev.csr_copy()
    dx, dy = dev.velocity_copy(with_dy=True)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    assert np.iscomplexobj(dx) and np.iscomplexobj(dy)
    assert np.max(np.abs(v - ov)) <= 1e-12 * np.max(np.abs(ov))
    assert np.max(np.abs(dx - odx)) <= 1e-12 * np.max(np.abs(odx))
    assert np.max(np.abs(dy - ody)) <= 1e-12 * np.max(np.abs(ody))
    with pytest.raises(L.Leo2dError) as e:
        dev.cond_moments(16, [0])                   # the conductivity recursion takes a real H
    assert e.value.code == L.ERR_ARG


@pytest.mark.parametrize("block", [0, 1, 4])
def test_cond_moments_match_oracle(dev, oracle_mod, block):
    """The deck's centre-cell targets and one more orbital, through the default
    block, blocks of 1 and blocks of 4 (a partial pass)."""
    text = W.tblg_flake_deck(3.0, 14, M=16)
    (orp, oci, ov, odx, _), tl = _oracle_setup(oracle_mod, text, 1)
    geom, _ = _device_setup(dev, text, L.PATTERN_PRUNED)
    targets = list(tl) + [7]
    M = 24
    want = oracle_mod.cond(orp, oci, ov, odx, M, targets)
    got = dev.cond_moments(M, targets, block=block)
    assert got.shape == (len(targets), M, M)
    scale = np.max(np.abs(want))
    assert np.max(np.abs(got - want)) <= 1e-12 * scale, np.max(np.abs(got - want)) / scale
    # targets per pass: the largest power of two within the request (default min(R, 16))
    want_block = block if block else 1 << (min(len(targets), 16).bit_length() - 1)
    assert dev.cond_timing().block == want_block


def test_cond_split_reduction(dev, oracle_mod):
    """A flake large enough (n > 3000) that the moment GEMM is split into row
    chunks whose partial products are summed afterwards."""
    text = W.tblg_flake_deck(2.0, 40, M=16)
    (orp, oci, ov, odx, _), tl = _oracle_setup(oracle_mod, text, 1)
    geom, _ = _device_setup(dev, text, L.PATTERN_PRUNED)
    assert geom.n > 3000
    M = 20
    want = oracle_mod.cond(orp, oci, ov, odx, M, list(tl))
    got = dev.cond_moments(M, list(tl))
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def test_cond_moments_multi_tile(dev, oracle_mod):
    """M = 160: two 128-wide output tiles per direction, the second partial, and
    the 3068-long reduction split into two row chunks (the second ending inside
    a 16-deep k tile)."""
    text = W.tblg_flake_deck(2.0, 36, M=16)
    (orp, oci, ov, odx, _), tl = _oracle_setup(oracle_mod, text, 1)
    geom, _ = _device_setup(dev, text, L.PATTERN_PRUNED)
    assert geom.n == 3068
    M = 160
    targets = list(tl)[:2]
    want = oracle_mod.cond(orp, oci, ov, odx, M, targets)
    got = dev.cond_moments(M, targets)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def test_cond_periodic_cell(dev, oracle_mod):
    text = W.tblg_cell_deck(3, 2, M=16)
    (orp, oci, ov, odx, _), tl = _oracle_setup(oracle_mod, text, 1)
    _device_setup(dev, text, L.PATTERN_PRUNED)
    M = 16
    want = oracle_mod.cond(orp, oci, ov, odx, M, list(tl))
    got = dev.cond_moments(M, list(tl))
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


@pytest.mark.parametrize("M", [16, 33, 128, 200])
def test_cond_transform_matches_oracle(dev, oracle_mod, M):
    rng = np.random.default_rng(M)
    T = 3 if M == 200 else 2
    Cm = rng.standard_normal((T, M, M))
    got = dev.cond_transform(Cm)
    for t in range(T):
        want = oracle_mod.cond_transform(Cm[t])
        assert np.max(np.abs(got[t] - want)) <= 1e-12 * np.max(np.abs(want)), (M, t)


def test_cond_error_paths(dev):
    text = W.tblg_flake_deck(3.0, 14, M=16)
    geom = L.Geometry(L.Deck(text))
    dev.set_geometry(geom)
    dev.pairs_build()
    job = geom.jobs()[0]
    dev.assemble(job, L.PATTERN_PRUNED)
    with pytest.raises(L.Leo2dError) as e:
        dev.cond_moments(16, [0])
    assert e.value.code == L.ERR_STATE
    dev.assemble_velocity(job, L.PATTERN_PRUNED)
    with pytest.raises(L.Leo2dError) as e:
        dev.cond_moments(16, [geom.n])
    assert e.value.code == L.ERR_ARG
