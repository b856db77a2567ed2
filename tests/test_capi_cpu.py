"""The C-ABI library without a GPU: it loads, exports every entry point that
include/dfm.h declares, and its host-only entry points (the FE pattern) match
the oracle; device entry points fail with a status instead of crashing."""
import ctypes as C

import numpy as np

from delfem_amd import _lib, meshgen


def test_exports_every_declared_symbol():
    L = _lib.lib()
    names = _lib.declared_symbols()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes table covers exactly the declared entry points
    assert sorted(_lib._SIGS) == names


def test_fem_pattern_matches_oracle(orc):
    for prob in (meshgen.poisson2d(9, seed=1), meshgen.solid3d(3, seed=1)):
        conn = np.ascontiguousarray(prob.conn, dtype=np.uint32)
        rowptr = np.zeros(prob.nnode + 1, dtype=np.uint32)
        n = C.c_size_t()
        _lib.call("dfm_fem_pattern", prob.nnode, conn.shape[0], conn.shape[1], conn.ctypes.data_as(C.c_void_p),
                  rowptr.ctypes.data_as(C.c_void_p), None, 0, C.byref(n))
        col = np.zeros(n.value, dtype=np.uint32)
        _lib.call("dfm_fem_pattern", prob.nnode, conn.shape[0], conn.shape[1], conn.ctypes.data_as(C.c_void_p),
                  rowptr.ctypes.data_as(C.c_void_p), col.ctypes.data_as(C.c_void_p), col.size, C.byref(n))
        want_ptr, want_col = orc.fem_pattern(prob.nnode, prob.conn)
        assert np.array_equal(rowptr, want_ptr) and np.array_equal(col, want_col)


def test_bad_arguments_report_errors():
    L = _lib.lib()
    conn = np.array([[0, 1, 7]], dtype=np.uint32)  # node 7 out of range for nnode = 3
    rc = L.dfm_fem_pattern(3, 1, 3, conn.ctypes.data_as(C.c_void_p), None, None, 0, None)
    assert rc == _lib.DFM_BAD_ARG
    assert b"nnode" in L.dfm_last_error()


def test_device_calls_without_device_fail_cleanly():
    import torch
    if torch.cuda.is_available():
        return
    L = _lib.lib()
    h = C.c_void_p()
    rc = L.dfm_bcrs_create(10, 3, C.byref(h))
    assert rc != _lib.DFM_OK and L.dfm_last_error()


def test_cpp_dropin_headers_compile(tmp_path):
    """include/delfem_amd/{matvec,ls}.hpp are valid C++17 on their own (the
    GPU test runs the program built from tests/cpp/shim_test.cpp)."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = os.path.join(root, "tests", "cpp", "shim_test.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(root, "include"),
                        src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
