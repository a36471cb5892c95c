"""The allocation-free KPM entry point (leo2d_kpm_moments_noalloc) against
leo2d_kpm_moments, through the C-ABI.

Bar: bit-exact.  Both entry points run the same kernels on the same matrix in
the reference row order (natural_order = 1) with the same grid and block, so
their moments agree bit for bit; only where the buffers live differs (the
library's static device scratch against fresh device allocations).

Two calls in a row, the second on a smaller matrix and a smaller block (the
scratch keeps the first call's contents), for every start form and scheme:
random phases (doubling), target starts (reference scheme; a partial last
block), caller vectors (reference scheme: the start-vector block and the upload
staging).  Sizes beyond the 64 MiB scratch and the mixed recursion are refused
with LEO2D_ERR_ARG before anything runs.
"""
import os

import numpy as np
import pytest

from conftest import ROOT

import leo2d_amd as L
from leo2d_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = L.Device(0)
    yield d
    d.close()


def _twin(dev):
    text = open(os.path.join(ROOT, "tests", "decks", "tblg_c0_twin.in")).read()
    geom = L.Geometry(L.Deck(text))
    dev.set_geometry(geom)
    dev.pairs_build()
    job = geom.jobs()[0]
    job.complex_matrix = 1
    job.k[0], job.k[1] = 0.013, -0.007
    dev.assemble(job, L.PATTERN_PRUNED)
    return geom.n


def _flake(dev):
    geom = L.Geometry(L.Deck(W.tblg_flake_deck(2.0, 21, M=64)))
    dev.set_geometry(geom)
    dev.pairs_build()
    job = geom.jobs()[0]
    job.energy_rescale, job.energy_shift = 10.5, 0.0
    dev.assemble(job, L.PATTERN_PRUNED)
    return geom.n


def _same(dev, **kw):
    ref = dev.kpm(natural_order=True, **kw)
    got = dev.kpm(noalloc=True, **kw)
    assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))
    return got


def test_noalloc_matches_kpm_moments(dev):
    # first call: the 8128-orbital twin cell, complex H(k), 64-vector blocks
    print("noalloc: c0 twin, R = 64 random phases, M = 64", flush=True)
    n = _twin(dev)
    assert n == 8128
    _same(dev, M=64, random_phase=(64, 3), scheme=L.SCHEME_DOUBLING)
    print("noalloc: c0 twin, 13 targets (block 8), M = 48", flush=True)
    _same(dev, M=48, targets=np.arange(13) * 601, block=8)
    print("noalloc: c0 twin, 40 complex vectors, M = 40", flush=True)
    rng = np.random.default_rng(5)
    V = rng.standard_normal((40, n)) + 1j * rng.standard_normal((40, n))
    _same(dev, M=40, vectors=V, block=64)

    # second call: a smaller matrix (real H) and smaller blocks
    print("noalloc: flake n = 1022, R = 16 random phases, M = 40", flush=True)
    n2 = _flake(dev)
    assert n2 == 1022
    _same(dev, M=40, random_phase=(16, 21), scheme=L.SCHEME_DOUBLING)
    print("noalloc: flake, 5 real vectors (block 8), M = 32", flush=True)
    V2 = np.random.default_rng(6).standard_normal((5, n2))
    _same(dev, M=32, vectors=V2, block=8)
    print("noalloc: flake, 3 targets, M = 36", flush=True)
    _same(dev, M=36, targets=[0, 511, 1021])


def test_noalloc_limits(dev):
    # a diagonal matrix of 300,000 rows: two 64-vector complex blocks are 614 MB
    print("noalloc: oversized workspace and mixed precision refused", flush=True)
    n = 300_000
    dev.csr_upload(np.arange(n + 1), np.arange(n), np.full(n, 0.5))
    with pytest.raises(L.Leo2dError) as e:
        dev.kpm(M=8, random_phase=(64, 1), noalloc=True)
    assert e.value.code == L.ERR_ARG
    assert "static scratch" in str(e.value)
    with pytest.raises(L.Leo2dError) as e:
        dev.kpm(M=8, random_phase=(4, 1), precision=L.PREC_MIXED, noalloc=True)
    assert e.value.code == L.ERR_ARG
    # a fitting size on the same matrix runs: mu_j = T_j(0.5) for every vector
    mu = dev.kpm(M=8, random_phase=(4, 1), noalloc=True, scheme=L.SCHEME_DOUBLING)
    ref = np.cos(np.arange(8) * np.arccos(0.5)) * n
    assert np.allclose(mu, ref[None, :], rtol=1e-12, atol=1e-9 * n)
