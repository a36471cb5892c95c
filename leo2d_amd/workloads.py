"""Decks and job preparation for commensurate twisted-bilayer-graphene cells.

These are the synthetic workloads the KPM path is measured and tested on:
an (m, n) commensurate tBLG cell (Locality::setupSupercell type 1,
src/locality.cpp:131-315) optionally tiled, with the GrapheneSimple coupling
model (nearest-neighbour t = -2.7 eV, t_perp = 0.48 exp(-(r - 3.35)/0.45) eV
for r < 6 A), a complex Bloch Hamiltonian H(k) and a Gershgorin rescale with
1 % padding.
"""
from __future__ import annotations

import math

import numpy as np

import leo2d_amd as L


def cell_orbitals(i: int) -> int:
    """Orbitals of the (i+1, i) commensurate cell: 4 (3 i^2 + 3 i + 1)."""
    return 4 * (3 * i * i + 3 * i + 1)


def tblg_cell_deck(m: int, n: int, tiles: int = 1, M: int = 1024, name: str = "tblg_cell",
                   target_sheets=(1,)) -> str:
    """Periodic two-sheet deck of the (m, n) cell tiled tiles x tiles
    (SUPERCELL_M_N = tiles*m tiles*n).  The orbital grid [-S, S) with
    S = 2 max(m, n) + min(m, n) + 1 (in tiled units) covers the cell with the
    centre cell (0, 0) inside it.  The targets are the centre-cell orbitals of
    target_sheets."""
    mm, nn = tiles * m, tiles * n
    S = 2 * max(mm, nn) + min(mm, nn) + 1
    ts = " ".join(str(t) for t in target_sheets)
    return f"""JOB_NAME = {name}
MAXSIZE = {S}
MINSIZE = -{S}
BOUNDARY_CONDITION = 1
SUPERCELL_M_N = {mm} {nn}
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
NUM_TARGET_SHEETS = {len(target_sheets)}
TARGET_SHEETS = {ts}
NUM_SHIFT_SHEETS = 1
SHIFT_SHEETS = 2
NSHIFTS = 1
ENERGY_RESCALE = 1
ENERGY_SHIFT = 0.0
POLY_ORDER = {M}
"""


def tblg_flake_deck(theta_deg: float, radius: float, M: int = 1024, name: str = "tblg_flake",
                    material: str = "GrapheneSimple", heterostrain=None, rescale: float = 10.5) -> str:
    """Open (disk) bilayer flake of the given radius in A (MAXSIZE also sets the
    disk radius, app/src/main.cpp:103), sheet 2 twisted by theta_deg;
    heterostrain = (eps, nu) strains sheet 2 uniaxially along x."""
    S = int(math.ceil(radius))
    hs = f"HETEROSTRAIN = {heterostrain[0]!r} {heterostrain[1]!r}\n" if heterostrain else ""
    return f"""JOB_NAME = {name}
MAXSIZE = {S}
MINSIZE = -{S}
BOUNDARY_CONDITION = 0
NUM_SHEETS = 2
START_SHEET 1
MATERIAL = {material}
HEIGHT = 0
ANGLE = 0
END_SHEET 1
START_SHEET 2
MATERIAL = {material}
HEIGHT = 3.35
ANGLE = {theta_deg!r}
END_SHEET 2
{hs}SOLVER_TYPE = SQ
NUM_TARGET_SHEETS = 1
TARGET_SHEETS = 1
NUM_SHIFT_SHEETS = 1
SHIFT_SHEETS = 2
NSHIFTS = 1
ENERGY_RESCALE = {rescale!r}
ENERGY_SHIFT = 0.0
POLY_ORDER = {M}
"""


def theta_mn(m: int, n: int) -> float:
    """Twist (degrees) of the (m, n) commensurate cell:
    cos theta = (n^2 + 4 n m + m^2) / (2 (n^2 + n m + m^2)), negative when
    m < n (Locality::setupSupercell type 1, src/locality.cpp:205-215)."""
    c = (n * n + 4 * n * m + m * m) / (2.0 * (n * n + n * m + m * m))
    t = math.degrees(math.acos(min(1.0, c)))
    return -t if m < n else t


# The small case: the (7, 6) cell (508 orbitals, 5.086 deg) four times over in
# each direction, 8,128 orbitals, M = 500, R = 16 random-phase vectors (seed 0),
# Gershgorin rescale with 1 % padding.  The reference builds open geometries as
# disks (MAX_R, app/src/main.cpp:116-120) and tiles a cell only periodically
# (BOUNDARY_CONDITION = 1), so the case exists in two reference-expressible forms.
C0_M, C0_R = 500, 16
C0_FLAKE_RADIUS = 58.54   # A: 8,108 orbitals, the disk count nearest 8,128 (2 pi r^2 x 0.3817 / A^2)


def c0_flake_deck(M: int = C0_M) -> str:
    """Open disk bilayer flake at the (7, 6) cell's twist whose orbital count
    is close to the 4 x 4 tiling's 8,128 (GrapheneSimple, real H)."""
    deck = tblg_flake_deck(theta_mn(7, 6), C0_FLAKE_RADIUS, M=M, name="tblg_c0_flake", rescale=1.0)
    S = int(math.ceil(2 * C0_FLAKE_RADIUS / 2.46)) + 1   # the index rhombus covers the disk
    deck = deck.replace(f"MAXSIZE = {int(math.ceil(C0_FLAKE_RADIUS))}\n",
                        f"MAXSIZE = {S}\nMAX_R = {C0_FLAKE_RADIUS!r}\n")
    return deck.replace(f"MINSIZE = -{int(math.ceil(C0_FLAKE_RADIUS))}\n", f"MINSIZE = -{S}\n")


def c0_twin_deck(M: int = C0_M) -> str:
    """The (7, 6) cell tiled 4 x 4 with periodic boundaries (SUPERCELL_M_N = 28 24)."""
    return tblg_cell_deck(7, 6, 4, M=M, name="tblg_c0_twin")


def reciprocal(geom) -> np.ndarray:
    """Reciprocal vectors b1, b2 (rows, 1/A) of the geometry's supercell."""
    a = np.array(geom.info.supercell)
    A3 = np.array([[a[0, 0], a[0, 1], 0.0], [a[1, 0], a[1, 1], 0.0], [0.0, 0.0, 1.0]])
    return (2 * np.pi * np.linalg.inv(A3).T)[:2, :2]


def assemble_rescaled(dev, geom, k_frac, pattern=L.PATTERN_PRUNED, pad: float = 1.01):
    """Fill H(k) for k = k_frac[0] b1 + k_frac[1] b2 (k_frac None: the real H
    of a flake or of a cell at Gamma) and rescale it into [-1, 1]:
    a = pad (hi - lo)/2, b = (hi + lo)/2 from the Gershgorin interval of the
    unscaled matrix.  Returns (job, a, b, nnz)."""
    job = geom.jobs()[0]
    if k_frac is None:
        job.complex_matrix = 0
    else:
        k = k_frac[0] * reciprocal(geom)[0] + k_frac[1] * reciprocal(geom)[1]
        job.complex_matrix = 1
        job.k[0], job.k[1] = float(k[0]), float(k[1])
    job.energy_rescale, job.energy_shift = 1.0, 0.0
    dev.assemble(job, pattern)
    lo, hi = dev.spectral_bounds()
    a = pad * (hi - lo) / 2.0
    b = (hi + lo) / 2.0
    job.energy_rescale, job.energy_shift = a, -b
    nnz = dev.assemble(job, pattern)
    return job, a, b, nnz
