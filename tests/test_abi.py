"""The C-ABI library loads without a GPU and exports every entry point the
public headers declare (no compute calls here)."""
import ctypes
import glob
import os
import re

import fastlbp_amd as fl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(flbp_\w+)\s*\(", re.M)


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        with open(h) as f:
            text = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
        names |= set(DECL.findall(text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) >= 28
    lib = ctypes.CDLL(fl.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_abi():
    from fastlbp_amd._native import SIGNATURES
    assert set(declared()) <= set(SIGNATURES)


def test_errors_without_gpu_are_status_codes():
    lib = fl.lib()
    assert lib.flbp_version().startswith(b"fastlbp")
    h = ctypes.c_void_p()
    rc = lib.flbp_graph_build(None, ctypes.byref(h))
    assert rc < 0 and b"bad arguments" in lib.flbp_last_error()


def test_size_limits_and_null_arguments_without_gpu():
    """Checks that come before any device call: sizes beyond the 32-bit
    index arrays, null handles and empty group lists are status codes."""
    from fastlbp_amd._native import CsrC, HistResultC
    lib = fl.lib()
    FLBP_EINVAL, FLBP_ERANGE = -1, -5
    big = CsrC()
    big.n_vars = (1 << 32) - 1
    h = ctypes.c_void_p()
    assert lib.flbp_plan_create(ctypes.byref(big), 0, ctypes.byref(h)) == FLBP_ERANGE
    assert b"2^32-1" in lib.flbp_last_error()
    big.n_vars, big.n_edges = 10, 1 << 32
    assert lib.flbp_plan_create(ctypes.byref(big), 0, ctypes.byref(h)) == FLBP_ERANGE
    assert lib.flbp_batch_create_groups(0, None, 0, None, ctypes.byref(h)) == FLBP_EINVAL
    res = HistResultC()
    assert lib.flbp_plan_run_hist_noalloc(None, 1e-6, 0, 10, 4, 1.0, 0.0, ctypes.byref(res), None) == FLBP_EINVAL
    assert lib.flbp_plan_run_hist(None, 1e-6, 0, 10, 4, 1.0, 0.0, ctypes.byref(res), None) == FLBP_EINVAL
