"""The C ABI without a GPU: symbol exports, status / message behaviour, loader errors -- CPU only.

Mirrors the reference's "pressure tests" (pcg/sw/tests/pcgtest.cpp:78-99) and the
exception -> status mapping of pcg/sw/src/pcg.cpp; every check here runs before
the library touches a device.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from hpc_amd import _lib
from hpc_amd.pcg import JPCG, Mode, Status, error_string

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def declared_functions():
    names = set()
    for h in os.listdir(INCLUDE):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b((?:xJPCG|xjpcg)_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding knows every one of them
    assert names <= set(_lib.declared_symbols())


def test_loader_archive_exports_pcg_api():
    out = subprocess.run(["nm", "-g", "--defined-only", _lib.LOADER_PATH], capture_output=True, text=True,
                         check=True).stdout
    for n in ["xJPCG_createHandle", "xJPCG_destroyHandle", "xJPCG_cscSymSolver", "xJPCG_cooSolver",
              "xJPCG_peekAtLastStatus", "xJPCG_getLastMessage", "xJPCG_getMetrics", "xJPCG_getErrorString"]:
        assert re.search(rf"\bT {n}\b", out), n


def test_error_strings():
    for s in Status:
        assert error_string(s) == f"XJPCG_STATUS_{s.name}"
    assert error_string(99) == "Unknown Error Code!"


def test_null_handle_paths():
    L = _lib.lib()
    it, res = ctypes.c_uint32(), ctypes.c_double()
    assert L.xJPCG_cooSolver(None, 4, 4, None, None, None, None, None, None, 10, 1e-8, ctypes.byref(it),
                             ctypes.byref(res), 0) == Status.NOT_INITIALIZED
    it64 = ctypes.c_int64()
    assert L.xJPCG_cscSymSolver(None, 4, 4, None, None, None, None, None, None, 10, 1e-8, ctypes.byref(it64),
                                ctypes.byref(res), 0) == Status.NOT_INITIALIZED
    assert L.xJPCG_peekAtLastStatus(None) == Status.NOT_INITIALIZED
    assert L.xJPCG_destroyHandle(None) == Status.NOT_INITIALIZED
    assert L.xJPCG_getMetrics(None, None) == Status.NOT_INITIALIZED
    assert L.xJPCG_getLastMessage(None).decode() == "Can't retrieve last error message from a null JPCG handle."
    assert L.xJPCG_createHandle(None, None) == Status.INVALID_VALUE


def small_coo():
    n = 4
    row = np.array([0, 0, 1, 1, 1, 2, 2, 2, 3, 3], np.uint32)
    col = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3], np.uint32)
    val = np.where(row == col, 4.0, -1.0)
    return n, row, col, val, np.full(n, 4.0), np.ones(n)


def test_pressure_sequence():
    """pcgtest.cpp:78-99 on a handle; without a GPU the handle reports why and still validates arguments."""
    n, row, col, val, diag, b = small_coo()
    with JPCG() as h:
        if has_gpu():
            assert h.create_status == Status.SUCCESS
        else:
            assert h.create_status == Status.INTERNAL_ERROR
            assert h.handle.value is not None
            assert "no HIP device" in h.last_message
        x = np.zeros(n)
        assert h.coo_solve_raw(0, row.size, row, col, val, diag, b, x, 10, 1e-8) == Status.INVALID_VALUE
        assert h.last_message == "INVALID VALUE ERROR: Wrong dimension size."
        assert h.last_status == Status.INVALID_VALUE
        assert h.coo_solve_raw(n, 0, row, col, val, diag, b, x, 10, 1e-8) == Status.INVALID_VALUE
        assert h.last_message == "INVALID VALUE ERROR: Wrong non-zero element size."
        st = h.coo_solve_raw(n, row.size, row, col, val, diag, b, x, 10, 1e-8, Mode.KEEP_NZ_LAYOUT)
        assert st == Status.INVALID_VALUE
        assert "wrong solver mode for the first call" in h.last_message
        assert h.coo_solve_raw(n, row.size, row, col, val, diag, b, x, 10, 1e-8, Mode.KEEP_MATRIX) == \
            Status.INVALID_VALUE
        assert h.coo_solve_raw(n, row.size, row, col, None, diag, b, x, 10, 1e-8) == Status.INVALID_VALUE
        assert h.last_message == "INVALID VALUE ERROR: Matrix is nullptr."
        assert h.coo_solve_raw(n, row.size, None, col, val, diag, b, x, 10, 1e-8) == Status.INVALID_VALUE
        # unsupported index base (mode 0x20): the reference's partitioner raises SpmNotSupported -> INTERNAL_ERROR
        assert h.coo_solve_raw(n, row.size, row, col, val, diag, b, x, 10, 1e-8, 0x20) == Status.INTERNAL_ERROR
        assert h.last_message.startswith("Matrix Operation ERROR")
        # out-of-range index
        bad = col.copy()
        bad[3] = 17
        assert h.coo_solve_raw(n, row.size, row, bad, val, diag, b, x, 10, 1e-8) == Status.INVALID_VALUE
        # Fortran base with a zero index underflows -> out of range
        assert h.coo_solve_raw(n, row.size, row, col, val, diag, b, x, 10, 1e-8, Mode.FORTRAN_INDEX) == \
            Status.INVALID_VALUE
        m = h.metrics()
        assert h.last_status == Status.SUCCESS and m.m_objInit >= 0.0


def test_large_coo_reports_first_bad_entry():
    """Index checks of a COO too large for one walking thread (split over host threads): the first
    out-of-range entry in input order is the one reported, as the serial walk of the reference loader stops
    there, and no device is needed to reach it."""
    n, nnz = 1000, 9_000_000
    rng = np.random.default_rng(4)
    row = rng.integers(0, n, nnz).astype(np.uint32)
    col = rng.integers(0, n, nnz).astype(np.uint32)
    val = np.ones(nnz)
    diag, b, x = np.ones(n), np.ones(n), np.zeros(n)
    with JPCG() as h:
        for bads in ([8_999_999], [6_100_000, 2_300_001], [4_500_000, 4_500_001, 1], [0, 8_000_000]):
            c = col.copy()
            for k in bads:
                c[k] = n + k % 7
            assert h.coo_solve_raw(n, nnz, row, c, val, diag, b, x, 10, 1e-8) == Status.INVALID_VALUE
            assert h.last_message == f"INVALID VALUE ERROR: matrix index out of range at entry {min(bads)}.", bads
        # Fortran base: a zero index underflows
        r1 = row + 1
        c1 = col + 1
        r1[7_777_777] = 0
        assert h.coo_solve_raw(n, nnz, r1, c1, val, diag, b, x, 10, 1e-8, Mode.FORTRAN_INDEX) == \
            Status.INVALID_VALUE
        assert h.last_message == "INVALID VALUE ERROR: matrix index out of range at entry 7777777."


def test_cscsym_argument_errors():
    n = 3
    with JPCG() as h:
        # column 1 lacks its diagonal: expansion count != 2 nnz - n -> reference internal error
        out = h.csc_sym_solve(n, [0, 1, 2, 2], [0, 2, 3, 4], [2.0, -1.0, 2.0, 2.0], np.full(n, 2.0),
                              np.ones(n), 10, 1e-8)
        if not has_gpu():
            assert out.status == Status.INTERNAL_ERROR
            assert "loadCscSym" in h.last_message or "CscSym" in h.last_message
        # Fortran mode with colPtr[0] < 1
        out = h.csc_sym_solve(n, [1, 2, 2, 3], [0, 2, 3, 4], [2.0, -1.0, 2.0, 2.0], np.full(n, 2.0), np.ones(n),
                              10, 1e-8, Mode.FORTRAN_INDEX)
        assert out.status == Status.INTERNAL_ERROR
        assert "colPtr[0] start from 1" in h.last_message


def test_loader_error_path():
    exe = os.path.join(ROOT, "tests", "c", "example_tridiag")
    env = dict(os.environ, XJPCG_LIBRARY="/nonexistent/libxjpcg.so")
    p = subprocess.run([exe, "--loader-error"], env=env, capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "LOADER 6 6 Dynamic loading error" in p.stdout
    assert "Cannot open library /nonexistent/libxjpcg.so" in p.stdout


def test_loader_reaches_library_without_gpu():
    if has_gpu():
        pytest.skip("checks the no-GPU message")
    exe = os.path.join(ROOT, "tests", "c", "example_tridiag")
    env = dict(os.environ, XJPCG_LIBRARY=_lib.LIB_PATH)
    p = subprocess.run([exe, "/tmp/xjpcg_cpu_probe"], env=env, capture_output=True, text=True, timeout=60)
    assert "CREATE 5 INTERNAL ERROR: no HIP device available." in p.stdout


def test_dev_api_validation_without_gpu():
    L = _lib.lib()
    op = ctypes.c_void_p()
    assert L.xjpcg_op_create(ctypes.byref(op), 0, 0, None, None, None, 0, None, None, None) == Status.INVALID_VALUE
    assert "Wrong dimension size" in L.xjpcg_dev_last_error().decode()
    assert L.xjpcg_op_destroy(None) == Status.NOT_INITIALIZED
    assert L.xjpcg_spmv(None, None, None, None) == Status.INVALID_VALUE  # null vectors checked first
    assert L.xjpcg_solve(None, None, None, None, 1, 1e-8, None, None, None, None, None) == Status.NOT_INITIALIZED


def test_reference_library_names():
    """libXilinxPcg.so / libXilinxPcgStatic.so (the reference's names) are this library."""
    for name in ("libXilinxPcg.so", "libXilinxPcgStatic.so"):
        path = os.path.join(ROOT, "hpc_amd", name)
        assert os.path.islink(path) and os.path.realpath(path) == os.path.realpath(_lib.LIB_PATH)
        L = ctypes.CDLL(path)
        assert hasattr(L, "xJPCG_cooSolver") and hasattr(L, "xJPCG_getErrorString")


REF_EXAMPLES = [os.path.join(ROOT, "oracle", "_ref", n) for n in ("ref_example_linked", "ref_example_loader")]


@pytest.mark.parametrize("exe", REF_EXAMPLES, ids=["linked", "loader"])
def test_reference_example_binaries_load_library_without_gpu(exe):
    """The reference's own examples/c/example.c, built against the reference pcg.h and linked with
    -lXilinxPcg or the reference's dlopen loader, reach this library; without a GPU, createHandle reports why."""
    if not os.path.exists(exe):
        pytest.skip("reference example not built (reference tree absent)")
    if has_gpu():
        pytest.skip("checks the no-GPU message")
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(ROOT, "hpc_amd"))
    p = subprocess.run([exe, "0"], env=env, capture_output=True, text=True, timeout=60)
    assert p.returncode == Status.INTERNAL_ERROR, p.stdout + p.stderr
    assert "INTERNAL ERROR: no HIP device available." in p.stderr


def test_step_api_errors():
    """xJPCG_setCooMat / setVec / run / getResults / getX: null handles, argument checks and call order."""
    L = _lib.lib()
    it, res = ctypes.c_uint32(), ctypes.c_double()
    assert L.xJPCG_setCooMat(None, 4, 4, None, None, None, 0) == Status.NOT_INITIALIZED
    assert L.xJPCG_setCscSymMat(None, 4, 4, None, None, None, 0) == Status.NOT_INITIALIZED
    assert L.xJPCG_setVec(None, 4, None, None) == Status.NOT_INITIALIZED
    assert L.xJPCG_run(None, 10, 1e-8, ctypes.byref(it), ctypes.byref(res)) == Status.NOT_INITIALIZED
    assert L.xJPCG_getResults(None, ctypes.byref(it), ctypes.byref(res)) == Status.NOT_INITIALIZED
    assert L.xJPCG_getX(None, None, 4) == Status.NOT_INITIALIZED
    n, row, col, val, diag, b = small_coo()
    with JPCG() as h:
        assert h.set_coo_mat(0, row, col, val) == Status.INVALID_VALUE
        assert h.last_message == "INVALID VALUE ERROR: Wrong dimension size."
        assert h.set_coo_mat(n, row, col, val, 0x03) == Status.INVALID_VALUE
        assert h.last_message == "INVALID VALUE ERROR: unknown solver mode."
        st, _, _ = h.run(10, 1e-8)
        assert st == Status.INVALID_VALUE and "before running the solver" in h.last_message
        st, _, _ = h.results()
        assert st == Status.INVALID_VALUE and "no solve has run" in h.last_message
        st, _ = h.get_x(n)
        assert st == Status.INVALID_VALUE and "no solve has run" in h.last_message
        if not has_gpu():
            assert h.set_coo_mat(n, row, col, val) == Status.INTERNAL_ERROR
            assert "no HIP device available" in h.last_message
