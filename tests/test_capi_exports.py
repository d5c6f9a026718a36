"""The C-ABI library loads and exports every entry point include/lsx.h declares (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lsx.h")
LIB = os.path.join(ROOT, "ls-stag-3d_parallel_amd", "liblsx.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lsx_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("lsx_create", "lsx_set_csr", "lsx_solve", "lsx_rebuild", "lsx_destroy", "lsx_precondition"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("liblsx.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_options_default_and_error_paths(lsx):
    """Host-only calls: defaults, argument checks and error reporting, no device work."""
    L = lsx.lib()
    o = lsx.Options()
    L.lsx_options_default(ctypes.byref(o))
    assert (o.krylov, o.restart, o.max_it, o.n_post) == (0, 13, 200, 3)
    assert abs(o.eps - 1e-5) < 1e-20 and abs(o.omega - 0.805) < 1e-15
    assert L.lsx_set_csr(None, 4, 4, 4, None, None, None, 0, 0) < 0
    assert L.lsx_solve(None, None, None, 0, None, None) < 0
    st = L.lsx_create(None, ctypes.byref(o), None)
    assert st < 0 and b"null" in L.lsx_last_error()
