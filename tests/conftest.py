import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


def golden_path(name: str) -> str:
    return os.path.join(GOLDEN, name)


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as orc

    orc.build()
    return orc


@pytest.fixture(scope="session")
def ref_lib():
    """The reference's own graphene.cpp / strained_graphene.cpp / sheet.cpp / spmatrix.cpp (oracle/_ref)."""
    import ctypes as C

    from oracle import oracle as orc

    path = orc.build_ref()
    if not path or not os.path.exists(path):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    # spmatrix.cpp references DMatrix (Eigen, not buildable here) from functions
    # this harness never calls: bind lazily so those symbols stay unresolved
    L = C.CDLL(path, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    L.ref_graphene_intra.restype = C.c_double
    L.ref_graphene_intra.argtypes = [C.c_int] * 4
    L.ref_graphene_inter.restype = C.c_double
    L.ref_graphene_inter.argtypes = [C.c_int, C.c_int] + [C.c_double] * 5
    L.ref_sheet_intra_pairs.restype = C.c_long
    L.ref_sheet_intra_pairs.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_void_p,
                                        C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_long, C.c_void_p]
    L.ref_sheet_positions.restype = C.c_int
    L.ref_sheet_positions.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_int]
    L.ref_set_sheet_material.restype = None
    L.ref_set_sheet_material.argtypes = [C.c_int]
    L.ref_strained_intra.restype = C.c_double
    L.ref_strained_intra.argtypes = [C.c_int] * 4
    L.ref_strained_intra_hop.restype = C.c_double
    L.ref_strained_intra_hop.argtypes = [C.c_int] * 4 + [C.c_double, C.c_double, C.c_void_p, C.c_double]
    L.ref_spmv.restype = None
    L.ref_spmv.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                           C.c_double, C.c_double]
    return L
