"""CPU tests of the C-ABI boundary: the library loads and exports every
function include/*.h declares; host-only entry points behave without a GPU."""
import ctypes as C
import glob
import os
import re
import subprocess

import pytest

from recon_scanner_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        for m in re.finditer(r"RECON_MVS_API\s+[\w\s\*]+?\b(\w+)\s*\(", txt):
            names.add(m.group(1))
    return sorted(names)


def exported(lib_path):
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_library_built():
    assert os.path.exists(N.HIP_LIB), "run __graft_entry__.build() first"


def test_every_declared_function_is_exported():
    names = declared_functions()
    assert len(names) >= 10
    ex = exported(N.HIP_LIB)
    missing = [n for n in names if n not in ex]
    assert not missing, f"declared but not exported: {missing}"


def test_library_loads_and_binds():
    lib = C.CDLL(N.HIP_LIB)
    for n in declared_functions():
        assert hasattr(lib, n)


def test_config_default_is_host_only_and_matches_mvs_h():
    lib = C.CDLL(N.HIP_LIB)
    lib.recon_mvs_config_default.argtypes = [C.POINTER(N.Config)]
    c = N.Config()
    lib.recon_mvs_config_default(C.byref(c))
    assert bytes(c) == bytes(N.config_default())


def test_struct_sizes_match_header():
    # layout of the POD structs (include/recon_mvs_hip.h)
    assert C.sizeof(N.Camera) == 8 * 9 + 4 * 2 + 8 * 12
    assert C.sizeof(N.NccResult) == 24
    assert C.sizeof(N.CostResult) == 32


def test_null_context_is_rejected():
    lib = C.CDLL(N.HIP_LIB)
    lib.recon_mvs_hip_synchronize.argtypes = [C.c_void_p]
    assert lib.recon_mvs_hip_synchronize(None) == N.ERR_INVALID_ARGUMENT
    lib.recon_mvs_hip_last_error.restype = C.c_char_p
    lib.recon_mvs_hip_last_error.argtypes = [C.c_void_p]
    assert lib.recon_mvs_hip_last_error(None) == b"null context"


def test_select_source_views_host_entry_matches_reference_golden():
    """recon_mvs_select_source_views is host-only (no GPU needed)."""
    from tests.helpers import cams_from_array, load
    g = load("select_views.npz")
    cams = cams_from_array(g["cams"])
    lib = C.CDLL(N.HIP_LIB)
    lib.recon_mvs_select_source_views.argtypes = [C.POINTER(N.Camera), C.c_int32,
                                                  C.POINTER(N.Camera), C.POINTER(N.Config),
                                                  C.c_void_p, C.c_int32]
    cfg = N.config_default()
    cfg.num_views = 3
    arr = (N.Camera * len(cams))(*cams)
    import numpy as np
    out = np.zeros(8, np.int32)
    n = lib.recon_mvs_select_source_views(C.byref(cams[0]), len(cams), arr, C.byref(cfg),
                                          out.ctypes.data, 8)
    assert out[:n].tolist() == g["selected"].tolist()


def test_cpp_caller_links_against_library():
    """A program written against include/recon/mvs.h builds and links."""
    binp = os.path.join(ROOT, "_build", "test_mvs_api")
    assert os.path.exists(binp)
    out = subprocess.run(["ldd", binp], capture_output=True, text=True).stdout
    assert "librecon_mvs_hip.so" in out and "not found" not in out
