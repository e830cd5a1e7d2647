"""The C-ABI library: it loads without a GPU, exports every function the
header declares, and its struct layouts match the ctypes mirror."""
import ctypes as C
import os
import re

from rbmd_amd import abi
from rbmd_amd.lib import load

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rbmd_mi355x.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rbmd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_functions():
    names = header_functions()
    assert "rbmd_md_step" in names and "rbmd_force_real" in names
    assert len(names) >= 30


def test_library_exports_every_header_function():
    lib = load()
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_sizes_match_ctypes():
    lib = load()
    lib.rbmd_abi_sizeof.restype = C.c_int32
    lib.rbmd_abi_sizeof.argtypes = [C.c_int32]
    assert lib.rbmd_abi_sizeof(0) == C.sizeof(abi.SysDesc)
    assert lib.rbmd_abi_sizeof(1) == C.sizeof(abi.Params)
    assert lib.rbmd_abi_sizeof(2) == C.sizeof(abi.Thermo)
    assert lib.rbmd_abi_sizeof(3) == -1


def test_version_and_error_strings():
    lib = load()
    lib.rbmd_version.restype = C.c_char_p
    lib.rbmd_last_error.restype = C.c_char_p
    assert lib.rbmd_version().decode()
    lib.rbmd_last_error()


def test_rbe_S_closed_form_without_device():
    """rbmd_rbe_S runs on the host: S = H^3 - 1 with the lattice sum H
    (RBEPSample.h:40-63) against a direct sum over a large grid."""
    import numpy as np

    import oracle

    lib = load()
    box = np.array([20.0, 20.0, 20.0])
    S = C.c_double()
    assert lib.rbmd_rbe_S(C.c_double(0.64), box.ctypes.data_as(C.POINTER(C.c_double)), C.byref(S)) == 0
    assert abs(S.value - oracle.rbe_S_direct(0.64, box, 60)) < 1e-12 * S.value


def test_bad_arguments_rejected_without_device():
    lib = load()
    lib.rbmd_last_error.restype = C.c_char_p
    S = C.c_double()
    box = (C.c_double * 3)(20.0, 20.0, 20.0)
    assert lib.rbmd_rbe_S(C.c_double(-1.0), box, C.byref(S)) == abi.RBMD_ERR_ARG
    assert b"bad" in lib.rbmd_last_error()
