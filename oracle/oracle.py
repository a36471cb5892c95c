"""CPU FP64 oracle for LEO2D's KPM hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module.  It is the checker, never the thing measured or shipped.

Pieces restated here (the numeric core is leo2d_oracle.c):
  * deck parsing          app/src/main.cpp:75-937 (token rules and quirks)
  * Job_params defaults   src/params/job_params.cpp:18-156
  * supercell setup       src/locality.cpp:131-400 (types 1 and 2)
  * relaxation field      src/geom/strain.cpp:258-311, 390-460 (STRAIN_TYPE = SUPERCELL)
  * job generation        src/locality.cpp:1222-1388 (SQ, LC), 1615-2249 (K_TYPE 0-5)
  * targets               src/geom/hstruct.cpp:1049-1080
  * .cheb writer          src/params/param_tools.cpp:19-285, 467-577
  * _matrix.dat writer    src/locality.cpp:4366-4383 (real), 5213-5228 (complex)
  * .kpts writer          src/params/param_tools.cpp:579-602
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

MAX_SHEETS = 8
MAT_GRAPHENE = 0
MAT_GRAPHENE_SIMPLE = 1
# src/materials/materials.cpp:15-21, plus the workload model name
MAT_GRAPHENE_SIMPLE_STRAIN = 2
MAT_STRAINED_GRAPHENE = 3
MATERIALS = {"Graphene": MAT_GRAPHENE, "StrainedGraphene": MAT_STRAINED_GRAPHENE, "GrapheneSimple": MAT_GRAPHENE_SIMPLE,
             "GrapheneSimpleStrain": MAT_GRAPHENE_SIMPLE_STRAIN}

GR_A = 2.4768
SQRT3_2 = 0.866025403784438646764
LATTICE = ((1.0 * GR_A, 0.0), (0.5 * GR_A, SQRT3_2 * GR_A))
# strained_graphene.h:50-52: sqrt(3)/2 rounded to 11 digits
STRAINED_LATTICE = ((1.0 * GR_A, 0.0), (0.5 * GR_A, 0.86602540378 * GR_A))


def lattice(mat: int):
    """Materials::lattice (materials.cpp:60-75)."""
    return STRAINED_LATTICE if mat == MAT_STRAINED_GRAPHENE else LATTICE


def build(force: bool = False) -> str:
    """Compile the oracle library (and oracle/_ref when the reference tree exists)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def build_ref() -> str | None:
    if not os.path.isdir("/root/reference/src"):
        return REF_PATH if os.path.exists(REF_PATH) else None
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)
    return REF_PATH


# --------------------------------------------------------------------------
# ctypes structures mirroring leo2d_oracle.h


class SheetDesc(C.Structure):
    _fields_ = [
        ("mat", C.c_int),
        ("min_shape", C.c_int * 2),
        ("max_shape", C.c_int * 2),
        ("max_R", C.c_double),
        ("boundary_condition", C.c_int),
        ("supercell", (C.c_double * 2) * 2),
        ("supercell_stride", (C.c_int * 2) * 2),
        ("angle", C.c_double),
        ("height", C.c_double),
        ("global_shifts_on", C.c_int),
        ("global_shift", C.c_double * 2),
    ]


class GeomDesc(C.Structure):
    _fields_ = [
        ("num_sheets", C.c_int),
        ("sheet", SheetDesc * MAX_SHEETS),
        ("boundary_condition", C.c_int),
        ("sc_search_size", C.c_int),
        ("supercell", (C.c_double * 2) * 2),
    ]


class Job(C.Structure):
    _fields_ = [
        ("shift", (C.c_double * 2) * MAX_SHEETS),
        ("energy_rescale", C.c_double),
        ("energy_shift", C.c_double),
        ("elec_on", C.c_int),
        ("E", C.c_double),
        ("mag_on", C.c_int),
        ("B", C.c_double),
        ("k", C.c_double * 2),
        ("complex_matrix", C.c_int),
        ("uniform_strain", C.c_int),
        ("u", (C.c_double * 2) * 2),
        ("pruned", C.c_int),
        ("sheet_strain_on", C.c_int),
        ("sheet_strain", ((C.c_double * 2) * 2) * MAX_SHEETS),
        ("strain_type", C.c_int),
        ("relax_max_k", C.c_int),
        ("relax", C.POINTER(C.c_double)),
        ("relax_b", (C.c_double * 2) * 2),
    ]


class Csr(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("nnz", C.c_int64),
        ("row_ptr", C.POINTER(C.c_int64)),
        ("col", C.POINTER(C.c_int32)),
        ("val", C.POINTER(C.c_double)),
        ("is_complex", C.c_int),
        ("val_dx", C.POINTER(C.c_double)),
        ("val_dy", C.POINTER(C.c_double)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.orc_geom_build.restype = vp
        L.orc_geom_build.argtypes = [C.POINTER(GeomDesc)]
        L.orc_geom_free.argtypes = [vp]
        L.orc_geom_n.argtypes = [vp]
        L.orc_geom_sheet_count.argtypes = [vp, C.c_int]
        L.orc_geom_arrays.argtypes = [vp, vp, vp]
        L.orc_grid_to_index.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_target_center.argtypes = [vp, C.c_int]
        L.orc_intralayer.restype = C.c_double
        L.orc_intralayer.argtypes = [C.c_int] * 5
        L.orc_interlayer.restype = C.c_double
        L.orc_interlayer.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_double] * 5
        L.orc_pairs_build.restype = vp
        L.orc_pairs_build.argtypes = [vp]
        L.orc_pairs_free.argtypes = [vp]
        L.orc_pairs_count.restype = C.c_int64
        L.orc_pairs_count.argtypes = [vp, C.c_int]
        L.orc_pairs_copy.argtypes = [vp, C.c_int, vp, vp, vp]
        L.orc_job_positions.argtypes = [vp, C.POINTER(Job), vp]
        L.orc_relax_disp.argtypes = [C.POINTER(Job), C.c_double, C.c_double, C.c_int, vp]
        L.orc_local_theta.argtypes = [vp, vp, vp]
        L.orc_assemble.argtypes = [vp, vp, C.POINTER(Job), C.POINTER(Csr)]
        L.orc_csr_free.argtypes = [C.POINTER(Csr)]
        L.orc_kpm.argtypes = [C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp,
                              C.c_int, C.c_int, C.c_int, vp]
        L.orc_spmv.argtypes = [C.c_int, vp, vp, vp, C.c_int, vp, vp, C.c_double, C.c_double]
        L.orc_phase_hash.restype = C.c_uint64
        L.orc_phase_hash.argtypes = [C.c_uint64] * 3
        L.orc_random_phase.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_int, vp]
        L.orc_jackson.argtypes = [C.c_int, vp]
        L.orc_energies.argtypes = [C.c_int, C.c_double, C.c_double, vp]
        L.orc_dos.argtypes = [C.c_int, vp, C.c_double, C.c_double, vp]
        L.orc_dos_trace.argtypes = [C.c_int, C.c_int, vp, C.c_double, C.c_double, C.c_double, vp]
        L.orc_cond.argtypes = [C.c_int, vp, vp, vp, vp, C.c_int, C.c_int, vp, vp]
        L.orc_cond_transform.argtypes = [C.c_int, vp, vp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# --------------------------------------------------------------------------
# deck parsing: app/src/main.cpp


def default_opts() -> dict:
    """Job_params defaults that affect this path (job_params.cpp:18-156)."""
    return {
        "job_name": "LEO2D_JOB", "solver_type": 0, "observable_type": 0, "solver_space": 0,
        "cond_transform": 1, "cond_poly_par": 0,
        "nShifts": 1, "num_shift_sheets": 1, "num_lc_points": 2,
        "lc_points": [[0.0, 0.0], [1.0, 1.0]], "num_sheets": 1, "boundary_condition": 0,
        "sc_search_size": 1, "global_shifts_on": 0, "strain_type": 0, "uniform_strain": 0,
        "matrix_save": 0, "matrix_pos_save": 0, "verbose_save": 1, "diagonalize": 0,
        "k_sampling": 0, "k_type": 0, "kpm_trace": 0, "kpm_trace_samps": -1,
        "energy_rescale": 20.0, "energy_shift": 0.0, "num_target_sheets": 1, "num_targets": 0,
        "poly_order": 20, "dos_transform": 1, "magOn": 0, "elecOn": 0, "B": 0.0, "E": 0.0,
        "chiral_on": 0, "ballistic_transport": 0, "mat_soc_on": 0, "mat_from_file": 0,
    }


@dataclass
class Sdata:
    mat: int
    min_shape: list
    max_shape: list
    max_R: float
    boundary_condition: int
    supercell: list = field(default_factory=lambda: [[0.0, 0.0], [0.0, 0.0]])
    stride: list = field(default_factory=lambda: [[0, 0], [0, 0]])


@dataclass
class Deck:
    opts: dict
    sdata: list
    heights: list
    angles: list


class DeckError(ValueError):
    pass


def parse_deck(text: str) -> Deck:
    """Restates the token loop of main.cpp:75-912: each line is split on
    single spaces; a recognised key consumes the following tokens ('=' and the
    value) and scanning continues with the rest of the line."""
    opts = default_opts()
    mins: list = []
    maxs: list = []
    max_R = 0.0
    bc = 0
    num_sheets = 0
    cur = -1
    sdata: list = []
    heights: list = []
    angles: list = []
    height = 0.0
    angle = 0.0

    def atoi(s):
        s = s.strip()
        m = 0
        sign = 1
        i = 0
        if i < len(s) and s[i] in "+-":
            sign = -1 if s[i] == "-" else 1
            i += 1
        while i < len(s) and s[i].isdigit():
            m = m * 10 + int(s[i])
            i += 1
        return sign * m

    def atof(s):
        s = s.strip()
        best = 0.0
        for j in range(len(s), 0, -1):
            try:
                best = float(s[:j])
                break
            except ValueError:
                continue
        return best

    for line in text.split("\n"):
        toks = line.split(" ")
        it = iter(toks)

        def nxt():
            return next(it, "")

        for tok in it:
            if tok == "JOB_NAME":
                nxt(); opts["job_name"] = nxt()
            elif tok == "MAXSIZE":
                nxt(); v = atoi(nxt()); maxs += [v, v, v]; max_R = float(v)
            elif tok == "MINSIZE":
                nxt(); v = atoi(nxt()); mins += [v, v, v]
            elif tok == "MAX_R":
                nxt(); max_R = atof(nxt())
            elif tok == "BOUNDARY_CONDITION":
                nxt(); bc = atoi(nxt()); opts["boundary_condition"] = bc
            elif tok == "SUPERCELL_GRID":
                nxt(); opts["x_supercell"] = atoi(nxt()); opts["y_supercell"] = atoi(nxt())
                opts["supercell_type"] = 2
            elif tok == "SUPERCELL_SEARCH_SIZE":
                nxt(); opts["sc_search_size"] = atoi(nxt())
            elif tok == "K_SAMPLING":
                nxt(); opts["k_sampling"] = atoi(nxt())
            elif tok == "K_TYPE":
                nxt(); opts["k_type"] = atoi(nxt())
            elif tok == "K_GRID":
                nxt(); opts["num_k1"] = atoi(nxt()); opts["num_k2"] = atoi(nxt())
            elif tok == "NUM_SHEETS":
                nxt(); num_sheets = atoi(nxt())
                sdata = [None] * num_sheets
                heights = [0.0] * num_sheets
                angles = [0.0] * num_sheets
                opts["num_sheets"] = num_sheets
            elif tok in ("CHIRAL_ON", "BALLISTIC_TRANSPORT", "MAT_SOC_ON", "MAT_FROM_FILE", "MLMC", "GSFE_Z_ON"):
                nxt(); opts[tok.lower()] = atoi(nxt())
            elif tok == "TRILAYER_SUPERCELL":
                # main.cpp:294-302: the supercell type changes only for the value 1
                nxt()
                if atoi(nxt()) == 1:
                    opts["supercell_type"] = 3
            elif tok == "H_SUPERCELL_MODIFY":
                nxt(); opts["h_supercell_modify"] = atoi(nxt())
            elif tok == "HEX_SUPERCELL_MODIFY":
                nxt(); opts["hex_supercell_modify"] = atoi(nxt())
            elif tok == "SUPERCELL_GROUPS":
                # main.cpp:131-155
                nxt()
                groups = [0] * num_sheets
                placed = 0
                for grp in range(2):
                    while placed < num_sheets:
                        v = nxt()
                        if v == "/":
                            break
                        sh = atoi(v)
                        if v == "" or not 1 <= sh <= num_sheets:
                            raise DeckError(f"SUPERCELL_GROUPS names no sheet {v!r}")
                        groups[sh - 1] = grp
                        placed += 1
                opts["sc_groups"] = groups
            elif tok == "SUPERCELL_M_N":
                nxt(); opts["m_supercell"] = atoi(nxt()); opts["n_supercell"] = atoi(nxt())
                opts["supercell_type"] = 1
            elif tok == "START_SHEET":
                cur = atoi(nxt()) - 1
                if cur + 1 > num_sheets:
                    raise DeckError(f"Cannot take in sheet {cur + 1} if NUM_SHEETS = {num_sheets}")
            elif tok == "END_SHEET":
                if cur == atoi(nxt()) - 1:
                    heights[cur] = height
                    angles[cur] = angle
                    cur = -1
            elif tok == "MATERIAL":
                nxt(); name = nxt()
                if name not in MATERIALS:
                    raise DeckError("Failed to find material! Aborting...")
                if not (0 <= cur < num_sheets):
                    raise DeckError("MATERIAL outside a START_SHEET/END_SHEET block")
                sdata[cur] = Sdata(MATERIALS[name], list(mins[:2]), list(maxs[:2]), max_R, bc)
            elif tok == "HEIGHT":
                nxt(); height = atof(nxt())
            elif tok == "ANGLE":
                nxt(); angle = (2 * math.pi * atof(nxt())) / (360.0)
            elif tok == "NSHIFTS":
                nxt(); opts["nShifts"] = atoi(nxt())
            elif tok == "ENERGY_RESCALE":
                nxt(); opts["energy_rescale"] = atof(nxt())
            elif tok == "ENERGY_SHIFT":
                nxt(); opts["energy_shift"] = atof(nxt())
            elif tok == "POLY_ORDER":
                nxt(); opts["poly_order"] = atoi(nxt())
            elif tok == "SOLVER_TYPE":
                nxt(); v = nxt()
                table = {"SQ": 1, "LC": 2, "MLMC_VAC": 3, "VD_FILE": 4, "STRAIN": 5, "FIXED_SHIFT": 6}
                if v in table:
                    opts["solver_type"] = table[v]
            elif tok == "NUM_LC_POINTS":
                nxt(); opts["num_lc_points"] = atoi(nxt())
            elif tok == "LC_POINTS":
                n = opts["num_lc_points"]
                nxt()
                pts = []
                for _ in range(n):
                    x = atof(nxt()); y = atof(nxt()); pts.append([x, y])
                opts["lc_points"] = pts
            elif tok == "OBSERVABLE_TYPE":
                nxt(); v = nxt()
                if v == "DOS":
                    opts["observable_type"] = 0
                elif v == "COND":
                    opts["observable_type"] = 1
            elif tok == "DOS_TRANSFORM":
                nxt(); opts["dos_transform"] = atoi(nxt())
            elif tok == "COND_TRANSFORM":
                nxt(); opts["cond_transform"] = atoi(nxt())
            elif tok == "COND_POLY_PAR":
                nxt(); opts["cond_poly_par"] = atoi(nxt())
            elif tok == "COND_POLY_PAR_SCALE":
                nxt(); opts["cond_poly_par_scale"] = atoi(nxt())
            elif tok == "SOLVER_SPACE":
                nxt(); v = nxt()
                if v[:1] == "R":
                    opts["solver_space"] = 0
                elif v[:1] == "M":
                    opts["solver_space"] = 1
            elif tok == "STRAIN_TYPE":
                # main.cpp:595-614: the value's first letter
                nxt(); v = nxt()
                codes = {"N": 0, "S": 1, "C": 2, "R": 3, "F": 4, "P": 5, "D": 6}
                if v[:1] in codes:
                    opts["strain_type"] = codes[v[:1]]
            elif tok == "STRAIN_FILE":
                nxt(); opts["strain_file"] = nxt()
            elif tok == "UNIFORM_STRAIN":
                nxt(); opts["uniform_strain"] = atoi(nxt())
            elif tok == "UNIFORM_STRAIN_MAT":
                nxt()
                u = [[0.0, 0.0], [0.0, 0.0]]
                for i in range(2):
                    for j in range(2):
                        u[i][j] = atof(nxt())
                opts["u_ij"] = u
            elif tok == "HETEROSTRAIN":
                nxt()
                eps = atof(nxt())
                nu = atof(nxt())
                opts["heterostrain"] = (eps, nu)
            elif tok == "MATRIX_SAVE":
                nxt(); opts["matrix_save"] = atoi(nxt())
            elif tok == "MATRIX_POS_SAVE":
                nxt(); opts["matrix_pos_save"] = atoi(nxt())
            elif tok == "VERBOSE_SAVE":
                nxt(); opts["verbose_save"] = atoi(nxt())
            elif tok == "DIAGONALIZE":
                nxt(); opts["diagonalize"] = atoi(nxt())
            elif tok == "KPM_TRACE":
                nxt(); opts["kpm_trace"] = atoi(nxt())
            elif tok == "KPM_TRACE_SAMPS":
                nxt(); opts["kpm_trace_samps"] = atoi(nxt())
            elif tok == "USE_B_FIELD":
                nxt(); opts["magOn"] = atoi(nxt())
            elif tok == "B_FIELD":
                nxt(); opts["B"] = atof(nxt())
            elif tok == "USE_E_FIELD":
                nxt(); opts["elecOn"] = atoi(nxt())
            elif tok == "E_FIELD":
                nxt(); opts["E"] = atof(nxt())
            elif tok == "NUM_TARGET_SHEETS":
                nxt(); opts["num_target_sheets"] = atoi(nxt())
            elif tok == "TARGET_SHEETS":
                nxt()
                opts["target_sheets"] = [atoi(nxt()) - 1 for _ in range(opts["num_target_sheets"])]
            elif tok == "NUM_SHIFT_SHEETS":
                nxt(); opts["num_shift_sheets"] = atoi(nxt())
            elif tok == "SHIFT_SHEETS":
                nxt()
                opts["shift_sheets"] = [atoi(nxt()) - 1 for _ in range(opts["num_shift_sheets"])]
            elif tok == "GLOBAL_SHIFTS_ON":
                nxt(); opts["global_shifts_on"] = atoi(nxt())
            elif tok == "GLOBAL_SHIFTS":
                gs = []
                for _ in range(opts["num_sheets"]):
                    nxt()
                    gs.append([atof(nxt()) for _ in range(3)])
                opts["global_shifts"] = gs
    # main.cpp:914-917
    if opts["poly_order"] % 4 != 0:
        raise DeckError(f"poly_order = {opts['poly_order']} is NOT divisible 4")
    if any(s is None for s in sdata):
        raise DeckError("a sheet has no MATERIAL")
    return Deck(opts=opts, sdata=sdata, heights=heights, angles=angles)


def setup_supercell(deck: Deck) -> None:
    """Locality::setupSupercell (locality.cpp:131-400) for types 1 and 2."""
    o = deck.opts
    if o.get("boundary_condition", 0) != 1:
        return
    typ = o.get("supercell_type")
    # Materials::lattice(sdata[0].mat): one unit cell for every sheet (locality.cpp:188-196)
    uc = lattice(deck.sdata[0].mat)
    if typ == 1:
        M = o["m_supercell"]
        N = o["n_supercell"]
        theta = math.acos((N * N + 4 * N * M + M * M) / (2.0 * (N * N + N * M + M * M)))
        if M < N:
            theta = -theta
        h = 1                          # locality.cpp:178-185
        if o.get("h_supercell_modify", 0) == 1:
            theta = theta + math.pi
            h = -1
        if len(deck.sdata) >= 3 and "sc_groups" not in o:
            raise DeckError("SUPERCELL_M_N with more than two sheets needs SUPERCELL_GROUPS")
        groups = [0, 1] if len(deck.sdata) < 3 else o["sc_groups"]
        for i, sd in enumerate(deck.sdata):
            g = groups[i]
            if g == 0:
                deck.angles[i] = 0
                A = ((N, M), (-M, M + N))
            else:
                deck.angles[i] = theta
                A = ((h * M, h * N), (-h * N, h * (M + N)))
            if o.get("hex_supercell_modify", 0) == 1:      # locality.cpp:222-233
                A = ((A[0][0] + A[0][1], A[0][1]), (A[1][0] + A[1][1], A[1][1]))
            sc = [[A[0][0] * uc[0][0] + A[0][1] * uc[1][0], A[0][0] * uc[0][1] + A[0][1] * uc[1][1]],
                  [A[1][0] * uc[0][0] + A[1][1] * uc[1][0], A[1][0] * uc[0][1] + A[1][1] * uc[1][1]]]
            sd.supercell = sc
            sd.stride = [[A[0][0], A[0][1]], [A[1][0], A[1][1]]]
            if i == 0:
                o["supercell"] = [list(sc[0]), list(sc[1])]
    elif typ == 2:
        X = o["x_supercell"]
        Y = o["y_supercell"]
        sc = [[X * uc[0][0], X * uc[0][1]], [Y * uc[1][0], Y * uc[1][1]]]
        for i, sd in enumerate(deck.sdata):
            th = deck.angles[i]
            c, s = math.cos(th), math.sin(th)
            sd.supercell = [[sc[0][0] * c + sc[0][1] * s, -sc[0][0] * s + sc[0][1] * c],
                            [sc[1][0] * c + sc[1][1] * s, -sc[1][0] * s + sc[1][1] * c]]
            sd.stride = [[X, 0], [0, Y]]
        o["supercell"] = sc
    else:
        raise DeckError("BOUNDARY_CONDITION = 1 needs SUPERCELL_M_N or SUPERCELL_GRID")


def geom_desc(deck: Deck) -> GeomDesc:
    d = GeomDesc()
    d.num_sheets = len(deck.sdata)
    if d.num_sheets > MAX_SHEETS:
        raise DeckError("too many sheets")
    o = deck.opts
    for k in ("chiral_on", "ballistic_transport", "mat_soc_on", "mlmc", "gsfe_z_on", "mat_from_file"):
        if o.get(k, 0) != 0:
            raise DeckError(f"{k} = 1 is outside this path")
    d.boundary_condition = o["boundary_condition"]
    d.sc_search_size = o["sc_search_size"]
    if d.boundary_condition == 1:
        for i in range(2):
            for j in range(2):
                d.supercell[i][j] = o["supercell"][i][j]
    for s, sd in enumerate(deck.sdata):
        sh = d.sheet[s]
        sh.mat = sd.mat
        sh.min_shape[0], sh.min_shape[1] = sd.min_shape
        sh.max_shape[0], sh.max_shape[1] = sd.max_shape
        sh.max_R = sd.max_R
        sh.boundary_condition = sd.boundary_condition
        for i in range(2):
            for j in range(2):
                sh.supercell[i][j] = sd.supercell[i][j]
                sh.supercell_stride[i][j] = sd.stride[i][j]
        sh.angle = deck.angles[s]
        sh.height = deck.heights[s]
        sh.global_shifts_on = o.get("global_shifts_on", 0)
        if sh.global_shifts_on == 1:
            sh.global_shift[0] = o["global_shifts"][s][0]
            sh.global_shift[1] = o["global_shifts"][s][1]
    return d


class Geometry:
    """Owner of an orc_geom (sheets, global index, positions)."""

    def __init__(self, desc: GeomDesc):
        self.desc = desc
        self.h = lib().orc_geom_build(C.byref(desc))
        if not self.h:
            raise DeckError("bad geometry")
        self.n = lib().orc_geom_n(self.h)
        self.index_to_grid = np.zeros(4 * self.n, dtype=np.int32)
        self.index_to_pos = np.zeros(3 * self.n, dtype=np.float64)
        lib().orc_geom_arrays(self.h, _p(self.index_to_grid), _p(self.index_to_pos))
        self._pairs = None

    def __del__(self):
        try:
            if self._pairs:
                lib().orc_pairs_free(self._pairs)
            lib().orc_geom_free(self.h)
        except Exception:
            pass

    def sheet_count(self, s):
        return lib().orc_geom_sheet_count(self.h, s)

    def target_center(self, s):
        return lib().orc_target_center(self.h, s)

    def pairs(self):
        if self._pairs is None:
            self._pairs = lib().orc_pairs_build(self.h)
        return self._pairs

    def pair_arrays(self, which: int):
        p = self.pairs()
        n = lib().orc_pairs_count(p, which)
        ij = np.zeros(2 * n, dtype=np.int32)
        sc = np.zeros(2 * n, dtype=np.int32)
        t = np.zeros(n, dtype=np.float64)
        lib().orc_pairs_copy(p, which, _p(ij), _p(sc), _p(t))
        return ij.reshape(-1, 2), sc.reshape(-1, 2), t

    def assemble(self, job: Job, with_dx: bool = False, with_dy: bool = False):
        """CSR of H (rp, ci, val); with_dx / with_dy also return the velocity
        matrices dxH / dyH on the same pattern, real or complex like H
        (locality.cpp:4182-4183, 4299-4300, 4939-4940, 5113-5114)."""
        csr = Csr()
        lib().orc_assemble(self.h, self.pairs(), C.byref(job), C.byref(csr))
        n, nnz = csr.n, csr.nnz
        rp = np.ctypeslib.as_array(csr.row_ptr, shape=(n + 1,)).copy()
        ci = np.ctypeslib.as_array(csr.col, shape=(max(nnz, 1),))[:nnz].copy()
        w = 2 if csr.is_complex else 1

        def take(ptr):
            a = np.ctypeslib.as_array(ptr, shape=(max(w * nnz, 1),))[: w * nnz].copy()
            return a.view(np.complex128) if csr.is_complex else a

        v = take(csr.val)
        dx = take(csr.val_dx)
        dy = take(csr.val_dy)
        lib().orc_csr_free(C.byref(csr))
        out = (rp, ci, v)
        if with_dx:
            out += (dx,)
        if with_dy:
            out += (dy,)
        return out

    def job_positions(self, job: Job):
        out = np.zeros(3 * self.n, dtype=np.float64)
        lib().orc_job_positions(self.h, C.byref(job), _p(out))
        return out

    def local_theta(self, i2pos: np.ndarray):
        out = np.zeros(self.n, dtype=np.float64)
        lib().orc_local_theta(self.h, _p(np.ascontiguousarray(i2pos, dtype=np.float64)), _p(out))
        return out


def target_list(deck: Deck, geom: Geometry) -> list:
    """Hstruct::getTargetList for solver types 1 and 2 (hstruct.cpp:1056-1079)."""
    out = []
    for s in deck.opts.get("target_sheets", [0])[: deck.opts["num_target_sheets"]]:
        c = geom.target_center(s)
        if c < 0:
            # the reference would index T_prev[-1] here; rejected instead
            raise DeckError("target sheet has no orbital in its centre cell")
        out += [c + orb for orb in range(2)]
    return out


def reciprocal(a):
    """Locality::getReciprocal(a) (locality.cpp:7215-7262)."""
    A = [[a[0][0], a[0][1], 0.0], [a[1][0], a[1][1], 0.0], [0.0, 0.0, 1.0]]

    def cross(x, y, d):
        if d == 0:
            return x[1] * y[2] - x[2] * y[1]
        if d == 1:
            return x[2] * y[0] - x[0] * y[2]
        return x[0] * y[1] - x[1] * y[0]

    den1 = 0.0
    den2 = 0.0
    for j in range(3):
        den1 += A[0][j] * cross(A[1], A[2], j)
        den2 += A[1][j] * cross(A[2], A[0], j)
    b = [[0.0, 0.0], [0.0, 0.0]]
    for k in range(2):
        b[0][k] = 2 * math.pi * cross(A[1], A[2], k) / den1
        b[1][k] = 2 * math.pi * cross(A[2], A[0], k) / den2
    return b


def sheet_reciprocal(deck: Deck, s: int):
    """Locality::getReciprocal(int s) (locality.cpp:7168-7213): sheet s's lattice
    rotated by its angle, strained by u_ij when BOUNDARY_CONDITION = 1 and
    UNIFORM_STRAIN = 1, then reciprocal(a)."""
    o = deck.opts
    th = deck.angles[s]
    lat = lattice(deck.sdata[s].mat)
    a = [[lat[i][0] * math.cos(th) - lat[i][1] * math.sin(th),
          lat[i][0] * math.sin(th) + lat[i][1] * math.cos(th)] for i in range(2)]
    if o["boundary_condition"] == 1 and o["uniform_strain"] == 1:
        u = o["u_ij"]
        a = [[(1.0 + u[0][0]) * t[0] + u[0][1] * t[1], u[1][0] * t[0] + (1.0 + u[1][1]) * t[1]] for t in a]
    return reciprocal(a)


def _fdiv(x: float, y: float) -> float:
    """x / y with IEEE semantics (inf / nan instead of ZeroDivisionError), as the
    reference's double division gives."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return float(np.float64(x) / np.float64(y))


def k_path(deck: Deck) -> list:
    """k vectors of one base job for K_TYPE 1-5 (locality.cpp:1655-2249), in job order.

    1: K -> Gamma -> M -> K of the supercell zone, c = 3 i / num_k1 (1655-1796)
    2: Gamma -> M -> K -> Gamma of sheet 1's zone (1799-1914)
    3: Gamma -> M -> K1 -> K2 -> Gamma of the supercell zone, c = 4 i / num_k1 (1916-1988)
    4: Gamma -> K -> M -> Gamma of the (M, N) cell, num_k1 points per line (1990-2163)
    5: K of sheet 1 -> K of sheet 2 -> K of sheet 3, 2 num_k1 + 1 points (2165-2249)"""
    o = deck.opts
    kt = o["k_type"]
    nk = o["num_k1"]
    ns = len(deck.sdata)
    pts = []
    if kt in (1, 3):
        if kt == 1 and ns < 2:
            raise DeckError("K_TYPE = 1 (twisted zone) needs two or more sheets")
        b = reciprocal(o["supercell"])
        g = (0.0, 0.0)
        k1 = ((1.0 / (2.0 * math.cos(math.pi / 6))) * (math.cos(-math.pi / 6) * b[1][0] - math.sin(-math.pi / 6) * b[1][1]),
              (1.0 / (2.0 * math.cos(math.pi / 6))) * (math.sin(-math.pi / 6) * b[1][0] + math.cos(-math.pi / 6) * b[1][1]))
        m = (b[1][0] / 2.0, b[1][1] / 2.0)
        if kt == 1:
            legs = (k1, g, m, k1)
            span = 3.0
        else:
            k2 = (math.cos(-math.pi / 3) * k1[0] - math.sin(-math.pi / 3) * k1[1],
                  math.sin(-math.pi / 3) * k1[0] + math.cos(-math.pi / 3) * k1[1])
            legs = (g, m, k1, k2, g)
            span = 4.0
        for i in range(nk):
            c = (span / float(nk)) * i
            seg = len(legs) - 2
            for s in range(len(legs) - 2):
                if c <= s + 1:
                    seg = s
                    break
            p, q = legs[seg], legs[seg + 1]
            hi = float(seg + 1)
            lo = float(seg)
            pts.append([(hi - c) * p[0] + (c - lo) * q[0], (hi - c) * p[1] + (c - lo) * q[1]])
    elif kt == 2:
        b1 = sheet_reciprocal(deck, 0)
        phi = math.pi / 3.0
        g = (0.0, 0.0)
        m = (b1[0][0] / 2.0, b1[0][1] / 2.0)
        k = ((m[0] * math.cos(phi / 2.0) - m[1] * math.sin(phi / 2.0)) / math.cos(phi / 2.0),
             (m[0] * math.sin(phi / 2.0) + m[1] * math.cos(phi / 2.0)) / math.cos(phi / 2.0))
        legs = (g, m, k, g)
        for i in range(nk):
            c = (3.0 / float(nk)) * i
            seg = 0 if c <= 1 else (1 if c <= 2 else 2)
            p, q = legs[seg], legs[seg + 1]
            pts.append([(seg + 1.0 - c) * p[0] + (c - seg) * q[0], (seg + 1.0 - c) * p[1] + (c - seg) * q[1]])
    elif kt == 4:
        if o.get("supercell_type") != 1:
            raise DeckError("K_TYPE = 4 needs SUPERCELL_M_N")
        M, N = o["m_supercell"], o["n_supercell"]
        if ns > 3:
            raise DeckError("K_TYPE = 4 takes at most three sheets")
        groups = [0, 1] if ns < 3 else o["sc_groups"]
        if len(groups) == 2:
            A = ((N, M), (-M, M + N)) if M > N else ((M, N), (-N, M + N))
        else:
            t = M * M + M * N + N * N
            A = ((0, t), (-t, t))
        uc = lattice(deck.sdata[0].mat)
        sc = [[A[0][0] * uc[0][0] + A[0][1] * uc[1][0], A[0][0] * uc[0][1] + A[0][1] * uc[1][1]],
              [A[1][0] * uc[0][0] + A[1][1] * uc[1][0], A[1][0] * uc[0][1] + A[1][1] * uc[1][1]]]
        bs = reciprocal(sc)
        r3 = math.sqrt(3)
        al = math.pi / 6
        rot = ((math.cos(al), math.sin(al)), (-math.sin(al), math.cos(al)))
        m = [0.5 * (bs[0][i] + bs[1][i]) for i in range(2)]
        k = [0.0, 0.0]
        for i in range(2):
            for j in range(2):
                k[i] = k[i] + rot[i][j] * bs[1][j] / r3
        g = (0.0, 0.0)
        s1 = _fdiv(k[1] - g[1], k[0] - g[0])
        s2 = _fdiv(m[1] - k[1], m[0] - k[0])
        s3 = _fdiv(g[1] - m[1], g[0] - m[0])
        b1 = _fdiv(k[0] * g[1] - g[0] * k[1], k[0] - g[0])
        b3 = _fdiv(g[0] * m[1] - m[0] * g[1], g[0] - m[0])
        l1 = k[0] - g[0]
        l2 = m[0] - k[0]
        l3 = g[0] - m[0]
        for i in range(3 * nk):
            if i < nk:
                pts.append([g[0] + i * l1 / nk, g[1] + s1 * i * l1 / nk + b1])
            elif i < 2 * nk:
                n = i - nk
                pts.append([k[0] + n * l2 / nk, k[1] + s2 * n * l2 / nk])
            else:
                n = i - 2 * nk
                if m[0] == m[1]:
                    ln = g[1] - m[1]
                    pts.append([m[0], m[1] + n * ln / nk])
                else:
                    pts.append([m[0] + n * l3 / nk, m[1] + s3 * n * l3 / nk + b3])
    elif kt == 5:
        if ns < 3:
            raise DeckError("K_TYPE = 5 needs three sheets")
        bl = [sheet_reciprocal(deck, s) for s in range(3)]
        K = [[1.0 / 3.0 * (2.0 * b[0][i] + b[1][i]) for i in range(2)] for b in bl]
        s1 = _fdiv(K[1][1] - K[0][1], K[1][0] - K[0][0])
        s2 = _fdiv(K[2][1] - K[1][1], K[2][0] - K[1][0])
        m1 = K[1][1] - s1 * K[1][0]
        m2 = K[2][1] - s2 * K[2][0]
        l1 = K[1][0] - K[0][0]
        l2 = K[2][0] - K[1][0]
        for i in range(2 * nk + 1):
            if i < nk:
                x = K[0][0] + i * l1 / nk
                pts.append([x, s1 * x + m1])
            elif i < 2 * nk:
                x = K[1][0] + (i - nk) * l2 / nk
                pts.append([x, s2 * x + m2])
            else:
                pts.append([K[2][0], K[2][1]])
    else:
        raise DeckError(f"K_TYPE {kt} is not a k-point scheme")
    return pts


def kpts_text(jobs: list) -> str:
    """Param_tools::saveKpts (param_tools.cpp:579-602) over the job list: fixed,
    10 decimals, a leading space on non-negative values, two tabs between the
    components and ' \\n' after the second."""
    out = []
    for job in jobs:
        if "k_vec" not in job:
            continue
        k0, k1 = job["k_vec"]
        out.append((" " if k0 >= 0 else "") + "%.10f" % k0 + "\t\t" + (" " if k1 >= 0 else "") + "%.10f" % k1 + " \n")
    return "".join(out)


def build_jobs(deck: Deck, targets: list) -> list:
    """rootChebSolve job list: SQ via recursiveShiftCalc (locality.cpp:1222-1269,
    1308-1327), LC (1330-1388), then the k-point expansion: the K_TYPE 0 grid
    (1615-1652) or the K_TYPE 1-5 paths of k_path (1655-2249)."""
    o = deck.opts
    ns = len(deck.sdata)
    jobs = []
    st = o["solver_type"]
    if st == 1:
        nS = o["nShifts"]
        nss = o["num_shift_sheets"]
        sss = o.get("shift_sheets", [])
        max_jobs = int(math.pow(nS * nS, nss))

        def rec(shifts, level):
            if level == 0:
                jobs.append({"shifts": [list(s) for s in shifts], "jobID": len(jobs) + 1,
                             "max_jobs": max_jobs, "target_list": list(targets)})
                return
            tar = sss[level - 1]
            for i in range(nS):
                for j in range(nS):
                    shifts[tar][0] = (1.0 / float(nS)) * i
                    shifts[tar][1] = (1.0 / float(nS)) * j
                    shifts[tar][2] = 0
                    rec(shifts, level - 1)

        rec([[0.0, 0.0, 0.0] for _ in range(ns)], nss)
    elif st == 6:
        # FIXED_SHIFT (locality.cpp:1580-1613)
        shifts = [[0.0, 0.0, 0.0] if s == 0 else [1.0 / 3.0, 1.0 / 3.0, 0] for s in range(ns)]
        jobs.append({"shifts": shifts, "jobID": 1, "max_jobs": 1, "target_list": list(targets)})
    elif st == 2:
        max_jobs = o["nShifts"]
        lc = o["lc_points"]
        nlc = o["num_lc_points"]
        for i in range(max_jobs):
            x = (1.0 / float(max_jobs)) * i
            total_x = x * (nlc - 1)
            idx = int(math.floor(total_x))
            eff = math.fmod(total_x, 1)
            shifts = [[0.0, 0.0, 0.0] for _ in range(ns)]
            shifts[ns - 1][0] = lc[idx][0] + eff * (lc[idx + 1][0] - lc[idx][0])
            shifts[ns - 1][1] = lc[idx][1] + eff * (lc[idx + 1][1] - lc[idx][1])
            shifts[ns - 1][2] = 0
            jobs.append({"shifts": shifts, "jobID": i + 1, "max_jobs": max_jobs,
                         "target_list": list(targets)})
    else:
        raise DeckError(f"solver_type {st} is outside this path")
    if o.get("k_sampling", 0) == 1 and o["boundary_condition"] != 1:
        raise DeckError("K_SAMPLING needs BOUNDARY_CONDITION = 1")
    if o.get("k_sampling", 0) == 1 and o.get("k_type", 0) != 0:
        pts = k_path(deck)
        kj = []
        for job in jobs:
            for kv in pts:
                nj = dict(job)
                nj["jobID"] = len(kj) + 1
                nj["k_vec"] = list(kv)
                kj.append(nj)
        jobs = kj
    elif o.get("k_sampling", 0) == 1:
        nk1, nk2 = o["num_k1"], o["num_k2"]
        b = reciprocal(o["supercell"])
        kj = []
        for job in jobs:
            for k1 in range(nk1):
                for k2 in range(nk2):
                    nj = dict(job)
                    nj["jobID"] = len(kj) + 1
                    nj["k_vec"] = [(k1 / float(nk1)) * b[0][0] + (k2 / float(nk2)) * b[1][0],
                                   (k1 / float(nk1)) * b[0][1] + (k2 / float(nk2)) * b[1][1]]
                    nj["k_vec_grid"] = [k1 / float(nk1), k2 / float(nk2)]
                    kj.append(nj)
        jobs = kj
    return jobs


def complex_matrix(opts: dict) -> int:
    """locality.cpp:2906-2908"""
    return int(opts["magOn"] == 1 or opts["solver_space"] == 1 or opts["k_sampling"] == 1
               or opts["chiral_on"] == 1 or opts["ballistic_transport"] == 1 or opts["mat_soc_on"] == 1)


def make_job(deck: Deck, job: dict, pruned: int = 0) -> Job:
    o = deck.opts
    j = Job()
    for s in range(len(deck.sdata)):
        j.shift[s][0] = job["shifts"][s][0]
        j.shift[s][1] = job["shifts"][s][1]
    j.energy_rescale = o["energy_rescale"]
    j.energy_shift = o["energy_shift"]
    j.elec_on = o["elecOn"]
    j.E = o["E"]
    j.mag_on = o["magOn"]
    j.B = o["B"]
    if "k_vec" in job:
        j.k[0], j.k[1] = job["k_vec"]
    j.complex_matrix = complex_matrix(o)
    j.uniform_strain = o["uniform_strain"]
    if j.uniform_strain == 1:
        for a in range(2):
            for b in range(2):
                j.u[a][b] = o["u_ij"][a][b]
    j.pruned = pruned
    if "heterostrain" in o:
        # uniaxial strain [[eps, 0], [0, -nu eps]] on the last sheet (open boundaries)
        if o["boundary_condition"] == 1:
            raise DeckError("HETEROSTRAIN needs open boundaries")
        eps, nu = o["heterostrain"]
        top = len(deck.sdata) - 1
        j.sheet_strain_on = 1
        j.sheet_strain[top][0][0] = eps
        j.sheet_strain[top][1][1] = -nu * eps
    st = o.get("strain_type", 0)
    if st == 1:
        if o["boundary_condition"] != 1:
            raise DeckError("STRAIN_TYPE = SUPERCELL needs a periodic supercell")
        max_k, coeffs = read_relaxation(o["strain_file"])
        set_relaxation(j, max_k, coeffs, reciprocal(o["supercell"]))
    elif st == 6:
        # DIRECT plane waves: StrainCalc::supercellDisp's built-in coefficients (strain.cpp:756-876)
        if o["boundary_condition"] != 1:
            raise DeckError("STRAIN_TYPE = DIRECT needs a periodic supercell")
        j.strain_type = 6
        b = reciprocal(o["supercell"])
        for a in range(2):
            for d in range(2):
                j.relax_b[a][d] = b[a][d]
    elif st != 0:
        raise DeckError(f"STRAIN_TYPE {st} is not on this path (NONE, SUPERCELL and DIRECT are)")
    return j


def read_relaxation(path: str):
    """StrainCalc::loadFourierConfigFile (strain.cpp:258-284): max_k, then the
    u_x, u_y and u_z coefficient rows of max_k (max_k + 1) / 2 + 1 values each,
    whitespace separated."""
    with open(path) as f:
        toks = f.read().split()
    max_k = int(toks[0])
    nc = max_k * (max_k + 1) // 2 + 1
    if len(toks) < 1 + 3 * nc:
        raise DeckError(f"relaxation file {path}: {3 * nc} coefficients expected")
    return max_k, np.array([float(t) for t in toks[1:1 + 3 * nc]]).reshape(3, nc)


def set_relaxation(j: Job, max_k: int, coeffs: np.ndarray, b) -> None:
    """Attach the Fourier relaxation field (STRAIN_TYPE = SUPERCELL) to a job."""
    c = np.ascontiguousarray(coeffs, dtype=np.float64).reshape(3, -1)
    j._relax_keep = c
    j.strain_type = 1
    j.relax_max_k = max_k
    j.relax = c.ctypes.data_as(C.POINTER(C.c_double))
    for a in range(2):
        for d in range(2):
            j.relax_b[a][d] = b[a][d]


def relax_disp(j: Job, x: float, y: float, s: int):
    d = np.zeros(3)
    lib().orc_relax_disp(C.byref(j), C.c_double(x), C.c_double(y), s, _p(d))
    return d


# --------------------------------------------------------------------------
# KPM and DoS


def kpm(rp, ci, val, M, R=None, targets=None, vecs=None, doubling_all=False, nthreads=1):
    """Faithful single-vector KPM per start state (computeDosKPM /
    computeDosTraceKPM).  Returns mu[R][M]."""
    L = lib()
    n = len(rp) - 1
    is_c = np.iscomplexobj(val)
    v = np.ascontiguousarray(val).view(np.float64) if is_c else np.ascontiguousarray(val, dtype=np.float64)
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    if targets is not None:
        t = np.ascontiguousarray(targets, dtype=np.int32)
        R = len(t)
        mu = np.zeros((R, M))
        rc = L.orc_kpm(n, _p(rp), _p(ci), _p(v), int(is_c), M, R, 0, _p(t), None, 0,
                       int(doubling_all), nthreads, _p(mu))
    else:
        vc = np.iscomplexobj(vecs)
        vv = np.ascontiguousarray(vecs)
        R = vv.shape[0]
        vflat = vv.view(np.float64) if vc else vv.astype(np.float64)
        mu = np.zeros((R, M))
        rc = L.orc_kpm(n, _p(rp), _p(ci), _p(v), int(is_c), M, R, 1, None, _p(vflat), int(vc),
                       int(doubling_all), nthreads, _p(mu))
    if rc != 0:
        raise ValueError("orc_kpm failed")
    return mu


def cond(rp, ci, val, val_dx, M, targets):
    """KPM conductivity moments C[t][n][m] (Locality::computeCondKPM,
    locality.cpp:6348-6727; real H, target starts), T_{M-1} of the beta set
    computed (the reference leaves it unassigned)."""
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    v = np.ascontiguousarray(val, dtype=np.float64)
    d = np.ascontiguousarray(val_dx, dtype=np.float64)
    t = np.ascontiguousarray(targets, dtype=np.int32)
    out = np.zeros((len(t), M, M))
    if lib().orc_cond(len(rp) - 1, _p(rp), _p(ci), _p(v), _p(d), M, len(t), _p(t), _p(out)) != 0:
        raise ValueError("orc_cond failed")
    return out


def cond_transform(Cm):
    """Param_tools::matrixResponseTransform (param_tools.cpp:838-905) of one
    M x M conductivity block."""
    Cm = np.ascontiguousarray(Cm, dtype=np.float64)
    M = Cm.shape[0]
    out = np.zeros((M, M))
    lib().orc_cond_transform(M, _p(Cm), _p(out))
    return out


def random_phase(n, R, seed, r0=0):
    v = np.zeros((R, n), dtype=np.complex128)
    lib().orc_random_phase(n, R, seed, r0, _p(v))
    return v


def jackson(M):
    g = np.zeros(M)
    lib().orc_jackson(M, _p(g))
    return g


def energies(M, a, b):
    E = np.zeros(M)
    lib().orc_energies(M, a, b, _p(E))
    return E


def dos(mu, a, b):
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    M = mu.shape[-1]
    out = np.zeros_like(mu)
    for idx in np.ndindex(mu.shape[:-1]):
        row = np.ascontiguousarray(mu[idx])
        o = np.zeros(M)
        lib().orc_dos(M, _p(row), a, b, _p(o))
        out[idx] = o
    return out


def dos_trace(mu, a, b, scale):
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    S, M = mu.shape
    out = np.zeros(M)
    lib().orc_dos_trace(M, S, _p(mu), a, b, scale, _p(out))
    return out


# --------------------------------------------------------------------------
# writers


def fmt_fixed(x: float) -> str:
    """std::fixed with the default precision 6."""
    return "%.6f" % x


def cheb_text(deck: Deck, results: list) -> str:
    """Param_tools::save + saveHeader for the DoS path (param_tools.cpp:19-80,
    228-285, 467-500, 564-577).  results: list of (job dict, dos[T][M])."""
    o = deck.opts
    M = o["poly_order"]
    a, b = o["energy_rescale"], o["energy_shift"]
    out = []
    E = energies(M, a, b)
    for job, rho in results:
        tl = job["target_list"]
        if o["verbose_save"] != 0:
            out.append("JOBID = %d \n" % job["jobID"])
            if o["solver_type"] in (1, 2):
                out.append("SHEET: SHIFT_X, SHIFT_Y, SHIFT_Z \n")
                for s in range(o["num_sheets"]):
                    sh = job["shifts"][s]
                    out.append("%d    : %s, %s, %s \n" % (s + 1, fmt_fixed(sh[0]), fmt_fixed(sh[1]), fmt_fixed(sh[2])))
                out.append("NUM_TAR = %d\n" % len(tl))
                if len(tl) != 0:
                    out.append("TAR_LIST: " + "".join("%d, " % t for t in tl[:-1]) + "%d \n" % tl[-1])
                else:
                    out.append("NO_TAR \n")
            if o["magOn"] == 1 or o["elecOn"] == 1:
                out.append("MAG_ON  = %d, B = %s \n" % (o["magOn"], fmt_fixed(o["B"])))
                out.append("ELEC_ON = %d, E = %s \n" % (o["elecOn"], fmt_fixed(o["E"])))
        if o["observable_type"] == 1:
            # conductivity (param_tools.cpp:287-317): only VERBOSE_SAVE = 0 writes the
            # matrices (the verbose block is commented out in the reference)
            if o["verbose_save"] == 0:
                if job["jobID"] == 1:
                    out.append("".join(fmt_fixed(e) + " " for e in E) + "\n\n")
                for t in range(len(tl)):
                    for row in rho[t]:
                        out.append("".join(fmt_fixed(x) + " " for x in row) + "\n")
                    out.append("\n")
            continue
        if o["verbose_save"] == 1:
            if job["jobID"] == 1:
                out.append("E: " + "".join(fmt_fixed(e) + " " for e in E) + "\n")
            out.append("T: \n")
            for t in range(len(tl)):
                row = rho[t]
                out.append("%d: " % tl[t] + "".join(fmt_fixed(x) + ", " for x in row[:-1]) + fmt_fixed(row[-1]) + "\n")
            out.append("\n")
        else:
            if job["jobID"] == 1:
                out.append("".join(fmt_fixed(e) + " " for e in E) + "\n")
            if o["kpm_trace"] == 1:
                out.append("".join(fmt_fixed(x) + " " for x in rho[:-1]) + fmt_fixed(rho[-1]) + "\n")
            else:
                for t in range(len(tl)):
                    row = rho[t]
                    out.append("".join(fmt_fixed(x) + " " for x in row[:-1]) + fmt_fixed(row[-1]) + "\n")
    return "".join(out)


def matrix_dat_text(rp, ci, val) -> str:
    """_matrix.dat: 1-based 'col, row, value' (real, |v| > 1e-12 only) or
    'col, row, re, im' (complex, every stored slot); default stream format."""
    out = []
    n = len(rp) - 1
    cpx = np.iscomplexobj(val)
    for i in range(n):
        for q in range(rp[i], rp[i + 1]):
            if cpx:
                out.append("%d, %d, %g, %g\n" % (ci[q] + 1, i + 1, val[q].real, val[q].imag))
            elif abs(val[q]) > 1e-12:
                out.append("%d, %d, %g\n" % (ci[q] + 1, i + 1, val[q]))
    return "".join(out)


def pos_dat_text(i2pos, i2g) -> str:
    """_pos.dat (locality.cpp:4425-4443): fixed, 12 decimals, 'x    y    z    sheet    orbital'."""
    p = np.asarray(i2pos).reshape(-1, 3)
    g = np.asarray(i2g).reshape(-1, 4)
    return "".join("%.12f    %.12f    %.12f    %d    %d\n" % (p[i, 0], p[i, 1], p[i, 2], g[i, 3], g[i, 2])
                   for i in range(len(p)))


def pair_pos_text(ij, i2pos) -> str:
    """_intra_pos.dat / _inter_pos.dat (locality.cpp:4446-4475): default stream
    format, 'x1, y1, z1, x2, y2, z2, i, j' per pair."""
    p = np.asarray(i2pos).reshape(-1, 3)
    return "".join("%g, %g, %g, %g, %g, %g, %d, %d\n" % (*p[a], *p[b], a, b) for a, b in np.asarray(ij).reshape(-1, 2))


# --------------------------------------------------------------------------
# KPM trace sampling: std::default_random_engine + std::normal_distribution as
# locality.cpp:2275-2276, 2301-2303 use them (libstdc++ 11, the image's
# library: bits/random.h, bits/random.tcc)


class MinstdRand0:
    """minstd_rand0 (= libstdc++'s default_random_engine): x <- 16807 x mod
    (2^31 - 1), default seed 1; returns the new state."""

    def __init__(self, seed: int = 1):
        self.x = seed % 2147483647 or 1

    def __call__(self) -> int:
        self.x = (16807 * self.x) % 2147483647
        return self.x


def generate_canonical(g) -> float:
    """std::generate_canonical<double, 53>: with range r = 2^31 - 2 it takes
    m = 2 draws (log2 r truncates to 30), sum = (g1 - 1) + (g2 - 1) r, / r^2."""
    r = 2147483646.0
    s, t = 0.0, 1.0
    for _ in range(2):
        s += float(g() - 1) * t
        t *= r
    v = s / t
    return v if v < 1.0 else math.nextafter(1.0, 0.0)


class NormalDistribution:
    """std::normal_distribution<double>(0, 1), Marsaglia polar method: each
    accepted pair (x, y) returns y * mult and keeps x * mult for the next call."""

    def __init__(self):
        self.saved = None

    def __call__(self, g) -> float:
        if self.saved is not None:
            v, self.saved = self.saved, None
            return v
        while True:
            x = 2.0 * generate_canonical(g) - 1.0
            y = 2.0 * generate_canonical(g) - 1.0
            r2 = x * x + y * y
            if not (r2 > 1.0 or r2 == 0.0):
                break
        mult = math.sqrt(-2 * math.log(r2) / r2)
        self.saved = x * mult
        return y * mult * 1.0 + 0.0


def trace_samples(opts: dict, n: int) -> int:
    """num_trace_samps (locality.cpp:2257-2270): KPM_TRACE_SAMPS, or for -1
    (int) 1e6 / n clamped to [1, 100]."""
    s = opts["kpm_trace_samps"]
    if s == -1:
        s = int(1000000 / float(n))
        s = min(max(s, 1), 100)
    return s


def run_deck(text: str, pruned: int = 0):
    """Whole DoS path for a deck on the CPU: returns (deck, geometry, jobs, results)."""
    deck = parse_deck(text)
    setup_supercell(deck)
    geom = Geometry(geom_desc(deck))
    targets = target_list(deck, geom)
    jobs = build_jobs(deck, targets)
    o = deck.opts
    results = []
    trace = o["kpm_trace"] == 1
    if trace:
        # one generator for all jobs, vectors drawn in job order (locality.cpp:2275-2310)
        gen, dist = MinstdRand0(), NormalDistribution()
        S = trace_samples(o, geom.n)
    cond_deck = o["observable_type"] == 1
    if cond_deck and (o["cond_poly_par"] != 0 or trace):
        raise DeckError("COND_POLY_PAR and KPM_TRACE are not on the conductivity path")
    for job in jobs:
        if cond_deck:
            # computeCondKPM + conductivityTransform (locality.cpp:3133-3147); real H only
            jb = make_job(deck, job, pruned)
            if jb.complex_matrix:
                raise DeckError("OBSERVABLE_TYPE = COND needs a real Hamiltonian")
            rp, ci, v, dx = geom.assemble(jb, with_dx=True)
            tl = job["target_list"]
            Cm = cond(rp, ci, v, dx, o["poly_order"], tl) if len(tl) else np.zeros((0, o["poly_order"], o["poly_order"]))
            if o["cond_transform"] == 1:
                Cm = np.array([cond_transform(c) for c in Cm])
            results.append((job, Cm))
            continue
        rp, ci, v = geom.assemble(make_job(deck, job, pruned))
        if trace:
            vecs = np.array([[dist(gen) for _ in range(geom.n)] for _ in range(S)])
            mu = kpm(rp, ci, v, o["poly_order"], vecs=vecs)
            rho = dos_trace(mu, o["energy_rescale"], o["energy_shift"], 1.0 / geom.n)
            if o["verbose_save"] == 1:
                # the verbose writer prints cheb_coeffs, which for a trace job holds the
                # samples' raw moments, one row per entry of the target list
                # (param_tools.cpp:39, 45-78; locality.cpp:3112-3130)
                if S < len(job["target_list"]):
                    raise DeckError("VERBOSE_SAVE = 1 with fewer trace samples than targets")
                rho = mu
        else:
            mu = kpm(rp, ci, v, o["poly_order"], targets=job["target_list"])
            rho = dos(mu, o["energy_rescale"], o["energy_shift"]) if o["dos_transform"] == 1 else mu
        results.append((job, rho))
    return deck, geom, jobs, results
