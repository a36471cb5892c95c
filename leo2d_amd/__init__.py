"""leo2d_amd -- MI355X-native drop-in for LEO2D's KPM hot path.

The product is libleo2d.so (HIP kernels + C-ABI, include/leo2d.h) and the
`leo2d` CLI.  This module is a thin ctypes binding of that C-ABI used by the
tests and the benchmark; it never falls back to a CPU path: if the built
library is missing it raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libleo2d.so")
BIN_PATH = os.path.join(HERE, "bin", "leo2d")

MAX_SHEETS = 8
MAX_TARGETS = 64

OK = 0
ERR_ARG = -1
ERR_DECK = -2
ERR_HIP = -3
ERR_ALLOC = -4
ERR_STATE = -5
ERR_SPECTRUM = -6
ERR_NAN = -7

PATTERN_REFERENCE = 0
PATTERN_PRUNED = 1
START_TARGETS = 0
START_VECTORS = 1
START_RANDOM_PHASE = 2
PREC_FP64 = 0
PREC_MIXED = 1
SCHEME_REFERENCE = 0
SCHEME_DOUBLING = 1
BATCH_ROW_GROUPS = 0
BATCH_ENTRY_SLOTS = 1
BATCH_ROW_CHUNKS = 2
BATCH_SELL = 3

# every symbol include/leo2d.h declares
EXPORTS = [
    "leo2d_deck_parse", "leo2d_deck_free", "leo2d_deck_get_int", "leo2d_deck_get_double",
    "leo2d_deck_get_string", "leo2d_geom_build", "leo2d_geom_free", "leo2d_geom_get_info",
    "leo2d_geom_arrays", "leo2d_jobs_count", "leo2d_job_get", "leo2d_create", "leo2d_destroy",
    "leo2d_last_error", "leo2d_set_geometry", "leo2d_pairs_build", "leo2d_pairs_method", "leo2d_pairs_copy",
    "leo2d_assemble", "leo2d_csr_upload", "leo2d_csr_info", "leo2d_csr_copy",
    "leo2d_spectral_bounds", "leo2d_kpm_moments", "leo2d_kpm_moments_noalloc", "leo2d_last_timing", "leo2d_dos",
    "leo2d_run_deck", "leo2d_batch_clear", "leo2d_batch_append", "leo2d_batch_info", "leo2d_batch_kpm",
    "leo2d_batch_last_timing", "leo2d_mlp_features", "leo2d_mlp_create", "leo2d_mlp_destroy",
    "leo2d_mlp_last_error", "leo2d_mlp_num_params", "leo2d_mlp_get_params", "leo2d_mlp_set_params",
    "leo2d_mlp_forward", "leo2d_mlp_gradients", "leo2d_mlp_train", "leo2d_mlp_train_ensemble", "leo2d_mlp_ensemble",
    "leo2d_assemble_velocity", "leo2d_velocity_copy", "leo2d_set_relaxation", "leo2d_read_relaxation", "leo2d_cond_moments", "leo2d_cond_transform",
    "leo2d_cond_last_timing",
]
MLP_FEATURES = 64
CONFIG_DIM = 6


class Leo2dError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"leo2d error {code}: {msg}")
        self.code = code


class GeomInfo(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("num_sheets", C.c_int32), ("sheet_count", C.c_int32 * MAX_SHEETS),
        ("boundary_condition", C.c_int32), ("supercell", (C.c_double * 2) * 2),
        ("angle", C.c_double * MAX_SHEETS),
    ]


class Job(C.Structure):
    _fields_ = [
        ("job_id", C.c_int32), ("max_jobs", C.c_int32), ("shift", (C.c_double * 3) * MAX_SHEETS),
        ("k_sampling", C.c_int32), ("k", C.c_double * 2), ("k_grid", C.c_double * 2),
        ("energy_rescale", C.c_double), ("energy_shift", C.c_double),
        ("elec_on", C.c_int32), ("mag_on", C.c_int32), ("E", C.c_double), ("B", C.c_double),
        ("complex_matrix", C.c_int32), ("uniform_strain", C.c_int32), ("u", (C.c_double * 2) * 2),
        ("num_targets", C.c_int32), ("targets", C.c_int32 * MAX_TARGETS),
        ("sheet_strain_on", C.c_int32), ("sheet_strain", ((C.c_double * 2) * 2) * MAX_SHEETS),
        ("strain_type", C.c_int32),
    ]


class KpmParams(C.Structure):
    _fields_ = [
        ("M", C.c_int32), ("R", C.c_int32), ("start", C.c_int32),
        ("targets", C.POINTER(C.c_int32)), ("vectors", C.POINTER(C.c_double)),
        ("vectors_complex", C.c_int32), ("seed", C.c_uint64), ("precision", C.c_int32),
        ("scheme", C.c_int32), ("block", C.c_int32), ("grid", C.c_int32), ("time_steps", C.c_int32),
        ("natural_order", C.c_int32),
    ]


class Timing(C.Structure):
    _fields_ = [
        ("recursion_ms", C.c_double), ("step_ms_avg", C.c_double), ("steps", C.c_int64),
        ("nnz", C.c_int64), ("n", C.c_int32), ("R", C.c_int32), ("M", C.c_int32), ("grid", C.c_int32),
        ("kernel", C.c_int32), ("reserved", C.c_int32),
    ]


class CondParams(C.Structure):
    _fields_ = [("M", C.c_int32), ("R", C.c_int32), ("targets", C.POINTER(C.c_int32)), ("block", C.c_int32)]


class CondTiming(C.Structure):
    _fields_ = [("recursion_ms", C.c_double), ("gemm_ms", C.c_double), ("block", C.c_int32), ("reserved", C.c_int32)]


class BatchParams(C.Structure):
    _fields_ = [
        ("M", C.c_int32), ("R", C.c_int32), ("start", C.c_int32), ("targets", C.POINTER(C.c_int32)),
        ("seed", C.c_uint64), ("precision", C.c_int32),
        ("tune", C.c_int32), ("graph", C.c_int32), ("mode", C.c_int32), ("tile_rows", C.c_int32),
        ("grid", C.c_int32),
    ]


class BatchTiming(C.Structure):
    _fields_ = [
        ("recursion_ms", C.c_double), ("tune_ms", C.c_double), ("mode", C.c_int32), ("tile_rows", C.c_int32),
        ("grid", C.c_int32), ("graph", C.c_int32), ("candidates", C.c_int32), ("jobs", C.c_int32),
        ("ntiles", C.c_int32), ("rows", C.c_int64), ("nnz", C.c_int64),
    ]


class MlpDesc(C.Structure):
    _fields_ = [("in_", C.c_int32), ("hidden", C.c_int32), ("layers", C.c_int32), ("out", C.c_int32),
                ("seed", C.c_uint64)]


class MlpTrainParams(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint64)]


_lib = None


def lib() -> C.CDLL:
    """Load the built HIP library (fails loudly when it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.leo2d_deck_parse.argtypes = [C.c_char_p, C.POINTER(vp), C.c_char_p, C.c_int]
        L.leo2d_deck_free.argtypes = [vp]
        L.leo2d_deck_free.restype = None
        L.leo2d_deck_get_int.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64)]
        L.leo2d_deck_get_double.argtypes = [vp, C.c_char_p, C.POINTER(C.c_double)]
        L.leo2d_deck_get_string.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_int]
        L.leo2d_geom_build.argtypes = [vp, C.POINTER(vp)]
        L.leo2d_geom_free.argtypes = [vp]
        L.leo2d_geom_free.restype = None
        L.leo2d_geom_get_info.argtypes = [vp, C.POINTER(GeomInfo)]
        L.leo2d_geom_arrays.argtypes = [vp, vp, vp]
        L.leo2d_jobs_count.argtypes = [vp, vp, C.POINTER(C.c_int32)]
        L.leo2d_job_get.argtypes = [vp, vp, C.c_int32, C.POINTER(Job)]
        L.leo2d_create.argtypes = [C.c_int32, C.POINTER(vp)]
        L.leo2d_destroy.argtypes = [vp]
        L.leo2d_destroy.restype = None
        L.leo2d_last_error.argtypes = [vp]
        L.leo2d_last_error.restype = C.c_char_p
        L.leo2d_set_geometry.argtypes = [vp, vp]
        L.leo2d_pairs_build.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.leo2d_pairs_method.argtypes = [vp, C.POINTER(C.c_int32)]
        L.leo2d_pairs_copy.argtypes = [vp, C.c_int32, vp, vp, vp]
        L.leo2d_assemble.argtypes = [vp, C.POINTER(Job), C.c_int32, C.POINTER(C.c_int64)]
        L.leo2d_csr_upload.argtypes = [vp, C.c_int32, C.c_int64, vp, vp, vp, C.c_int32]
        L.leo2d_csr_info.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
        L.leo2d_csr_copy.argtypes = [vp, vp, vp, vp]
        L.leo2d_assemble_velocity.argtypes = [vp, C.POINTER(Job), C.c_int32, C.POINTER(C.c_int64)]
        L.leo2d_velocity_copy.argtypes = [vp, vp, vp]
        L.leo2d_set_relaxation.argtypes = [vp, C.c_int32, vp]
        L.leo2d_read_relaxation.argtypes = [C.c_char_p, C.POINTER(C.c_int32), vp, C.POINTER(C.c_int64)]
        L.leo2d_cond_moments.argtypes = [vp, C.POINTER(CondParams), vp]
        L.leo2d_cond_transform.argtypes = [vp, C.c_int32, C.c_int32, vp, vp]
        L.leo2d_cond_last_timing.argtypes = [vp, C.POINTER(CondTiming)]
        L.leo2d_spectral_bounds.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.leo2d_kpm_moments.argtypes = [vp, C.POINTER(KpmParams), vp]
        L.leo2d_kpm_moments_noalloc.argtypes = [vp, C.POINTER(KpmParams), vp]
        L.leo2d_last_timing.argtypes = [vp, C.POINTER(Timing)]
        L.leo2d_dos.argtypes = [vp, vp, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, vp, vp]
        L.leo2d_run_deck.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_int32]
        L.leo2d_mlp_features.argtypes = [vp, C.c_int64, vp]
        L.leo2d_mlp_create.argtypes = [C.c_int32, C.POINTER(MlpDesc), C.POINTER(vp)]
        L.leo2d_mlp_destroy.argtypes = [vp]
        L.leo2d_mlp_destroy.restype = None
        L.leo2d_mlp_last_error.argtypes = [vp]
        L.leo2d_mlp_last_error.restype = C.c_char_p
        L.leo2d_mlp_num_params.argtypes = [vp]
        L.leo2d_mlp_num_params.restype = C.c_int64
        L.leo2d_mlp_get_params.argtypes = [vp, vp]
        L.leo2d_mlp_set_params.argtypes = [vp, vp]
        L.leo2d_mlp_forward.argtypes = [vp, vp, C.c_int64, vp]
        L.leo2d_mlp_gradients.argtypes = [vp, vp, vp, C.c_int32, C.POINTER(C.c_double), vp]
        L.leo2d_mlp_train.argtypes = [vp, vp, vp, C.c_int64, C.POINTER(MlpTrainParams), vp]
        L.leo2d_mlp_train_ensemble.argtypes = [vp, C.c_int32, vp, vp, C.c_int64, C.POINTER(MlpTrainParams), vp, vp]
        L.leo2d_mlp_ensemble.argtypes = [C.POINTER(vp), C.c_int32, vp, C.c_int64, C.c_int32, vp, vp]
        L.leo2d_batch_clear.argtypes = [vp]
        L.leo2d_batch_append.argtypes = [vp, C.POINTER(C.c_int32)]
        L.leo2d_batch_info.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.leo2d_batch_kpm.argtypes = [vp, C.POINTER(BatchParams), vp]
        L.leo2d_batch_last_timing.argtypes = [vp, C.POINTER(BatchTiming)]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Deck:
    """Parsed hstruct deck (app/src/main.cpp semantics), owned by the library."""

    def __init__(self, text: str):
        self.text = text
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = lib().leo2d_deck_parse(text.encode(), C.byref(h), err, 512)
        if rc != OK:
            raise Leo2dError(rc, err.value.decode())
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().leo2d_deck_free(self.h)
            self.h = None

    def get_int(self, key):
        v = C.c_int64()
        if lib().leo2d_deck_get_int(self.h, key.encode(), C.byref(v)) != OK:
            raise KeyError(key)
        return v.value

    def get_double(self, key):
        v = C.c_double()
        if lib().leo2d_deck_get_double(self.h, key.encode(), C.byref(v)) != OK:
            raise KeyError(key)
        return v.value

    def get_string(self, key):
        b = C.create_string_buffer(1024)
        if lib().leo2d_deck_get_string(self.h, key.encode(), b, 1024) != OK:
            raise KeyError(key)
        return b.value.decode()


class Geometry:
    def __init__(self, deck: Deck):
        self.deck = deck
        h = C.c_void_p()
        rc = lib().leo2d_geom_build(deck.h, C.byref(h))
        if rc != OK:
            raise Leo2dError(rc, "geometry build failed")
        self.h = h
        self.info = GeomInfo()
        lib().leo2d_geom_get_info(self.h, C.byref(self.info))
        self.n = self.info.n

    def __del__(self):
        if getattr(self, "h", None):
            lib().leo2d_geom_free(self.h)
            self.h = None

    def arrays(self):
        i2g = np.zeros(4 * self.n, dtype=np.int32)
        i2p = np.zeros(3 * self.n, dtype=np.float64)
        lib().leo2d_geom_arrays(self.h, _p(i2g), _p(i2p))
        return i2g, i2p

    def jobs(self):
        cnt = C.c_int32()
        rc = lib().leo2d_jobs_count(self.deck.h, self.h, C.byref(cnt))
        if rc != OK:
            raise Leo2dError(rc, "job list failed")
        out = []
        for q in range(cnt.value):
            j = Job()
            lib().leo2d_job_get(self.deck.h, self.h, q, C.byref(j))
            out.append(j)
        return out


class Device:
    """One leo2d_handle: a HIP stream and the device state of one worker."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib().leo2d_create(device, C.byref(h))
        if rc != OK:
            raise Leo2dError(rc, "leo2d_create failed (no GPU?)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().leo2d_destroy(self.h)
            self.h = None

    __del__ = close

    def _check(self, rc):
        if rc != OK:
            raise Leo2dError(rc, lib().leo2d_last_error(self.h).decode())

    def set_geometry(self, geom: Geometry):
        self._check(lib().leo2d_set_geometry(self.h, geom.h))

    def pairs_build(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(lib().leo2d_pairs_build(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def pairs_method(self) -> int:
        """1 when the last pairs_build staged each row's lists in LDS (one
        wavefront per row), 0 for the thread-per-row kernels."""
        v = C.c_int32()
        self._check(lib().leo2d_pairs_method(self.h, C.byref(v)))
        return v.value

    def pairs_copy(self, which: int, count: int):
        ij = np.zeros(2 * max(count, 1), dtype=np.int32)
        sc = np.zeros(2 * max(count, 1), dtype=np.int32)
        t = np.zeros(max(count, 1), dtype=np.float64)
        self._check(lib().leo2d_pairs_copy(self.h, which, _p(ij), _p(sc), _p(t) if which == 0 else None))
        return ij[: 2 * count].reshape(-1, 2), sc[: 2 * count].reshape(-1, 2), t[:count]

    def assemble(self, job: Job, pattern: int = PATTERN_REFERENCE) -> int:
        nnz = C.c_int64()
        self._check(lib().leo2d_assemble(self.h, C.byref(job), pattern, C.byref(nnz)))
        return nnz.value

    def assemble_velocity(self, job: Job, pattern: int = PATTERN_REFERENCE) -> int:
        """H and the velocity matrices dxH, dyH on its pattern (real or complex like H)."""
        nnz = C.c_int64()
        self._check(lib().leo2d_assemble_velocity(self.h, C.byref(job), pattern, C.byref(nnz)))
        return nnz.value

    def set_relaxation(self, max_k: int, coeffs=None):
        """Fourier relaxation field for jobs with strain_type = 1 (coeffs [3][nc]); max_k < 0 clears it."""
        c = None if coeffs is None else np.ascontiguousarray(coeffs, dtype=np.float64)
        self._check(lib().leo2d_set_relaxation(self.h, max_k, _p(c) if c is not None else None))

    def velocity_copy(self, with_dy: bool = False):
        """dxH (and dyH) values; complex128 arrays for a complex H."""
        _, nnz, cpx = self.csr_info()
        w = 2 if cpx else 1
        dx = np.zeros(w * max(nnz, 1), dtype=np.float64)
        dy = np.zeros(w * max(nnz, 1), dtype=np.float64)
        self._check(lib().leo2d_velocity_copy(self.h, _p(dx), _p(dy) if with_dy else None))
        dx, dy = dx[: w * nnz], dy[: w * nnz]
        if cpx:
            dx, dy = dx.view(np.complex128), dy.view(np.complex128)
        return (dx, dy) if with_dy else dx

    def cond_moments(self, M, targets, block=0):
        """KPM conductivity moments C[t][n][m] (computeCondKPM)."""
        t = np.ascontiguousarray(targets, dtype=np.int32)
        p = CondParams()
        p.M, p.R, p.block = M, len(t), block
        p.targets = t.ctypes.data_as(C.POINTER(C.c_int32))
        out = np.zeros((len(t), M, M))
        self._check(lib().leo2d_cond_moments(self.h, C.byref(p), _p(out)))
        return out

    def cond_transform(self, Cm):
        """2D Jackson / DCT-III transform of conductivity blocks [T][M][M]."""
        Cm = np.ascontiguousarray(Cm, dtype=np.float64)
        if Cm.ndim == 2:
            Cm = Cm[None]
        out = np.zeros_like(Cm)
        self._check(lib().leo2d_cond_transform(self.h, Cm.shape[1], Cm.shape[0], _p(Cm), _p(out)))
        return out

    def cond_timing(self) -> CondTiming:
        t = CondTiming()
        self._check(lib().leo2d_cond_last_timing(self.h, C.byref(t)))
        return t

    def csr_upload(self, rp, ci, val):
        rp = np.ascontiguousarray(rp, dtype=np.int64)
        ci = np.ascontiguousarray(ci, dtype=np.int32)
        cpx = np.iscomplexobj(val)
        v = np.ascontiguousarray(val, dtype=np.complex128 if cpx else np.float64)
        self._check(lib().leo2d_csr_upload(self.h, len(rp) - 1, len(ci), _p(rp), _p(ci),
                                           _p(v.view(np.float64)), int(cpx)))

    def csr_info(self):
        n, nnz, cpx = C.c_int32(), C.c_int64(), C.c_int32()
        self._check(lib().leo2d_csr_info(self.h, C.byref(n), C.byref(nnz), C.byref(cpx)))
        return n.value, nnz.value, cpx.value

    def csr_copy(self):
        n, nnz, cpx = self.csr_info()
        rp = np.zeros(n + 1, dtype=np.int64)
        ci = np.zeros(max(nnz, 1), dtype=np.int32)
        v = np.zeros(max(nnz, 1) * (2 if cpx else 1), dtype=np.float64)
        self._check(lib().leo2d_csr_copy(self.h, _p(rp), _p(ci), _p(v)))
        ci = ci[:nnz]
        v = v[: nnz * (2 if cpx else 1)]
        return rp, ci, (v.view(np.complex128) if cpx else v)

    def spectral_bounds(self):
        lo, hi = C.c_double(), C.c_double()
        self._check(lib().leo2d_spectral_bounds(self.h, C.byref(lo), C.byref(hi)))
        return lo.value, hi.value

    def kpm(self, M, targets=None, vectors=None, random_phase=None, precision=PREC_FP64,
            scheme=SCHEME_REFERENCE, block=0, grid=0, time_steps=False, natural_order=False, noalloc=False):
        """Chebyshev moments mu[R][M] on the device Hamiltonian (noalloc: through
        leo2d_kpm_moments_noalloc, FP64 in the reference row order, workspace in the
        library's static device scratch)."""
        p = KpmParams()
        p.M = M
        p.precision = precision
        p.scheme = scheme
        p.block = block
        p.grid = grid
        p.time_steps = int(time_steps)
        p.natural_order = int(natural_order)
        keep = []
        if targets is not None:
            t = np.ascontiguousarray(targets, dtype=np.int32)
            keep.append(t)
            p.R = len(t)
            p.start = START_TARGETS
            p.targets = t.ctypes.data_as(C.POINTER(C.c_int32))
        elif vectors is not None:
            cpx = np.iscomplexobj(vectors)
            v = np.ascontiguousarray(vectors, dtype=np.complex128 if cpx else np.float64)
            keep.append(v)
            p.R = v.shape[0]
            p.start = START_VECTORS
            p.vectors_complex = int(cpx)
            p.vectors = v.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))
        else:
            R, seed = random_phase
            p.R = R
            p.start = START_RANDOM_PHASE
            p.seed = seed
        mu = np.zeros((p.R, M))
        entry = lib().leo2d_kpm_moments_noalloc if noalloc else lib().leo2d_kpm_moments
        self.last_rc = entry(self.h, C.byref(p), _p(mu))
        self._check(self.last_rc)
        return mu

    def timing(self) -> Timing:
        t = Timing()
        lib().leo2d_last_timing(self.h, C.byref(t))
        return t

    # ---- batched jobs ----
    def batch_clear(self):
        self._check(lib().leo2d_batch_clear(self.h))

    def batch_append(self) -> int:
        """Append the current device Hamiltonian as the next job of the batch."""
        j = C.c_int32()
        self._check(lib().leo2d_batch_append(self.h, C.byref(j)))
        return j.value

    def batch_info(self):
        jobs, rows, nnz = C.c_int32(), C.c_int64(), C.c_int64()
        self._check(lib().leo2d_batch_info(self.h, C.byref(jobs), C.byref(rows), C.byref(nnz)))
        return jobs.value, rows.value, nnz.value

    def batch_kpm(self, M, R, seed=0, precision=PREC_FP64, tune=False, graph=True, mode=BATCH_ROW_GROUPS,
                  tile_rows=0, grid=0, targets=None):
        """Moments mu[jobs][R][M] of every job in the batch: random-phase starts,
        or unit vectors on targets[jobs][R] (orbital indices inside each job)."""
        p = BatchParams()
        p.M, p.R, p.seed, p.precision = M, R, seed, precision
        p.start = START_RANDOM_PHASE
        if targets is not None:
            tg = np.ascontiguousarray(targets, dtype=np.int32)
            p.start = START_TARGETS
            p.targets = tg.ctypes.data_as(C.POINTER(C.c_int32))
        p.tune, p.graph, p.mode, p.tile_rows, p.grid = int(tune), int(graph), mode, tile_rows, grid
        jobs = self.batch_info()[0]
        mu = np.zeros((jobs, R, M))
        self.last_rc = lib().leo2d_batch_kpm(self.h, C.byref(p), _p(mu))
        self._check(self.last_rc)
        return mu

    def batch_timing(self) -> BatchTiming:
        t = BatchTiming()
        lib().leo2d_batch_last_timing(self.h, C.byref(t))
        return t

    def dos(self, mu, a, b, scale=1.0):
        mu = np.ascontiguousarray(np.atleast_2d(mu), dtype=np.float64)
        T, M = mu.shape
        rho = np.zeros_like(mu)
        E = np.zeros(M)
        self._check(lib().leo2d_dos(self.h, _p(mu), T, M, a, b, scale, _p(E), _p(rho)))
        return E, rho

    def run_deck(self, text: str, out_prefix: str, pattern: int = PATTERN_REFERENCE):
        self._check(lib().leo2d_run_deck(self.h, text.encode(), out_prefix.encode(), pattern))


# ---- learned surrogate ------------------------------------------------------
def read_relaxation(path: str):
    """max_k and the [3][nc] coefficients of a STRAIN_FILE, read by the library."""
    mk, cnt = C.c_int32(), C.c_int64()
    rc = lib().leo2d_read_relaxation(path.encode(), C.byref(mk), None, C.byref(cnt))
    if rc != OK:
        raise Leo2dError(rc, "cannot read " + path)
    c = np.zeros(cnt.value, dtype=np.float64)
    rc = lib().leo2d_read_relaxation(path.encode(), C.byref(mk), _p(c), C.byref(cnt))
    if rc != OK:
        raise Leo2dError(rc, "short relaxation file " + path)
    return mk.value, c.reshape(3, -1)


def mlp_features(cfg) -> np.ndarray:
    """[n][64] Fourier features of configurations [n][6] = (b1, b2, exx, eyy, exy, theta_deg)."""
    c = np.ascontiguousarray(np.atleast_2d(cfg), dtype=np.float64)
    f = np.zeros((c.shape[0], MLP_FEATURES), dtype=np.float32)
    rc = lib().leo2d_mlp_features(_p(c), c.shape[0], _p(f))
    if rc != OK:
        raise Leo2dError(rc, "leo2d_mlp_features: invalid configuration")
    return f


class Mlp:
    """One surrogate network on a device (BF16 MFMA training and inference)."""

    def __init__(self, device=0, in_=MLP_FEATURES, hidden=256, layers=3, out=1024, seed=0):
        d = MlpDesc()
        d.in_, d.hidden, d.layers, d.out, d.seed = in_, hidden, layers, out, seed
        h = C.c_void_p()
        rc = lib().leo2d_mlp_create(device, C.byref(d), C.byref(h))
        if rc != OK:
            raise Leo2dError(rc, "leo2d_mlp_create failed")
        self.h = h
        self.dims = [in_] + [hidden] * layers + [out]

    def close(self):
        if getattr(self, "h", None):
            lib().leo2d_mlp_destroy(self.h)
            self.h = None

    __del__ = close

    def _check(self, rc):
        if rc != OK:
            raise Leo2dError(rc, lib().leo2d_mlp_last_error(self.h).decode())

    @property
    def num_params(self) -> int:
        return lib().leo2d_mlp_num_params(self.h)

    def params(self) -> np.ndarray:
        p = np.zeros(self.num_params, dtype=np.float32)
        self._check(lib().leo2d_mlp_get_params(self.h, _p(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params
        self._check(lib().leo2d_mlp_set_params(self.h, _p(p)))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.dims[-1]), dtype=np.float32)
        self._check(lib().leo2d_mlp_forward(self.h, _p(x), x.shape[0], _p(y)))
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.num_params, dtype=np.float32)
        loss = C.c_double()
        self._check(lib().leo2d_mlp_gradients(self.h, _p(x), _p(t), x.shape[0], C.byref(loss), _p(g)))
        return loss.value, g

    def train(self, x, t, epochs=10, batch=256, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, seed=0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        p = MlpTrainParams()
        p.epochs, p.batch, p.lr, p.beta1, p.beta2, p.eps, p.seed = epochs, batch, lr, beta1, beta2, eps, seed
        losses = np.zeros(max(epochs, 1))
        self._check(lib().leo2d_mlp_train(self.h, _p(x), _p(t), x.shape[0], C.byref(p), _p(losses)))
        return losses[:epochs]


def mlp_train_ensemble(nets, x, t, seeds, epochs=10, batch=256, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """Train the networks together (leo2d_mlp_train_ensemble): network k shuffles
    with seeds[k]; returns the per-epoch losses [len(nets)][epochs]."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    p = MlpTrainParams()
    p.epochs, p.batch, p.lr, p.beta1, p.beta2, p.eps, p.seed = epochs, batch, lr, beta1, beta2, eps, 0
    arr = (C.c_void_p * len(nets))(*[m.h for m in nets])
    sd = np.ascontiguousarray(seeds, dtype=np.uint64)
    losses = np.zeros((len(nets), max(epochs, 1)))
    rc = lib().leo2d_mlp_train_ensemble(arr, len(nets), _p(x), _p(t), x.shape[0], C.byref(p), _p(sd), _p(losses))
    if rc != OK:
        raise Leo2dError(rc, lib().leo2d_mlp_last_error(nets[0].h).decode())
    return losses[:, :epochs]


def mlp_ensemble(nets, x, with_mean=True, blocks=1):
    """(mean [n][out] or None, spread [n] (blocks = 1) or [n][blocks]) of an
    ensemble of networks of one shape; spread = RMS over each block of the
    outputs of the ensemble standard deviation."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    arr = (C.c_void_p * len(nets))(*[m.h for m in nets])
    mean = np.zeros((x.shape[0], nets[0].dims[-1]), dtype=np.float32) if with_mean else None
    spread = np.zeros((x.shape[0], blocks), dtype=np.float32)
    rc = lib().leo2d_mlp_ensemble(arr, len(nets), _p(x), x.shape[0], blocks, _p(mean), _p(spread))
    if rc != OK:
        raise Leo2dError(rc, lib().leo2d_mlp_last_error(nets[0].h).decode())
    return mean, (spread[:, 0] if blocks == 1 else spread)
