"""Python mirror of the pcg.h handle API (reference pcg/sw/include/pcg.h).

`JPCG` wraps one XJPCG_Handle_t of libxjpcg.so.  Its methods take host numpy
arrays and keep the reference's argument meaning, status codes and error
behaviour: a failing call returns its status (it does not raise), and
`last_status` / `last_message` report it, exactly as the C API does.
"""
from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass

import numpy as np

from ._lib import lib, ptr


class Status(enum.IntEnum):
    """XJPCG_Status_t (reference pcg.h:41-49)."""
    SUCCESS = 0
    NOT_INITIALIZED = 1
    ALLOC_FAILED = 2
    INVALID_VALUE = 3
    EXECUTION_FAILED = 4
    INTERNAL_ERROR = 5
    DYNAMIC_LOADING_ERROR = 6


class Mode(enum.IntFlag):
    """XJPCG_Mode_t (reference pcg.h:79-85)."""
    DEFAULT = 0x00
    KEEP_NZ_LAYOUT = 0x01
    KEEP_MATRIX = 0x02
    C_INDEX = 0x00
    FORTRAN_INDEX = 0x10


class Metric(ctypes.Structure):
    """XJPCG_Metric_t (reference pcg.h:62-67), seconds."""
    _fields_ = [("m_objInit", ctypes.c_double), ("m_matProc", ctypes.c_double),
                ("m_vecProc", ctypes.c_double), ("m_solver", ctypes.c_double)]


RECORD_DTYPE = np.dtype([("id", "<u4"), ("iter", "<u4"), ("vec_size", "<u4"), ("max_iter", "<u4"),
                         ("tols", "<f8"), ("res", "<f8"), ("rz", "<f8"), ("alpha", "<f8"), ("beta", "<f8"),
                         ("clock", "<u8")])
assert RECORD_DTYPE.itemsize == 64


@dataclass
class SolveOutput:
    status: Status
    x: np.ndarray
    iterations: int
    residual: float


def error_string(code: int) -> str:
    return lib().xJPCG_getErrorString(int(code)).decode()


def _arr(a, dtype):
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=dtype)


class JPCG:
    """One solver session (XJPCG_Handle_t)."""

    def __init__(self, device: int | str | None = None):
        self._L = lib()
        h = ctypes.c_void_p()
        path = None if device is None else str(device).encode()
        self.create_status = Status(self._L.xJPCG_createHandle(ctypes.byref(h), path))
        self._h = h

    # -- lifecycle -----------------------------------------------------------
    def close(self):
        if self._h:
            self._L.xJPCG_destroyHandle(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # -- status --------------------------------------------------------------
    @property
    def last_status(self) -> Status:
        return Status(self._L.xJPCG_peekAtLastStatus(self._h))

    @property
    def last_message(self) -> str:
        return self._L.xJPCG_getLastMessage(self._h).decode(errors="replace")

    def metrics(self) -> Metric:
        m = Metric()
        self._L.xJPCG_getMetrics(self._h, ctypes.byref(m))
        return m

    def history(self) -> np.ndarray:
        """Per-iteration records of the last solve (record 0 = set-up)."""
        count = ctypes.c_uint64()
        self._L.xJPCG_getHistory(self._h, None, 0, ctypes.byref(count))
        out = np.zeros(count.value, dtype=RECORD_DTYPE)
        if count.value:
            self._L.xJPCG_getHistory(self._h, out.ctypes.data, count.value, ctypes.byref(count))
        return out

    # -- solves --------------------------------------------------------------
    def coo_solve(self, n, row, col, data, diag, b, max_iter, tol, mode=Mode.DEFAULT, x=None) -> SolveOutput:
        """xJPCG_cooSolver; row/col may be None for KEEP_NZ_LAYOUT / KEEP_MATRIX."""
        row = _arr(row, np.uint32)
        col = _arr(col, np.uint32)
        data = _arr(data, np.float64)
        diag = _arr(diag, np.float64)
        b = _arr(b, np.float64)
        if x is None:
            x = np.zeros(int(n), dtype=np.float64)
        it = ctypes.c_uint32(0)
        res = ctypes.c_double(0.0)
        st = self._L.xJPCG_cooSolver(self._h, int(n), 0 if data is None else int(data.size), ptr(row), ptr(col),
                                     ptr(data), ptr(diag), ptr(b), ptr(x), int(max_iter), float(tol),
                                     ctypes.byref(it), ctypes.byref(res), int(mode))
        return SolveOutput(Status(st), x, it.value, res.value)

    def coo_solve_raw(self, n, nnz, row, col, data, diag, b, x, max_iter, tol, mode=Mode.DEFAULT) -> Status:
        """xJPCG_cooSolver with every argument passed through as given (None = NULL)."""
        it = ctypes.c_uint32(0)
        res = ctypes.c_double(0.0)
        return Status(self._L.xJPCG_cooSolver(self._h, int(n), int(nnz), ptr(row), ptr(col), ptr(data), ptr(diag),
                                              ptr(b), ptr(x), int(max_iter), float(tol), ctypes.byref(it),
                                              ctypes.byref(res), int(mode)))

    def csc_sym_solve(self, n, row_idx, col_ptr, data, diag, b, max_iter, tol, mode=Mode.DEFAULT,
                      x=None) -> SolveOutput:
        """xJPCG_cscSymSolver (half-stored symmetric CSC, diagonal included)."""
        row_idx = _arr(row_idx, np.int64)
        col_ptr = _arr(col_ptr, np.int64)
        data = _arr(data, np.float64)
        diag = _arr(diag, np.float64)
        b = _arr(b, np.float64)
        if x is None:
            x = np.zeros(int(n), dtype=np.float64)
        it = ctypes.c_int64(0)
        res = ctypes.c_double(0.0)
        st = self._L.xJPCG_cscSymSolver(self._h, int(n), int(data.size), ptr(row_idx), ptr(col_ptr), ptr(data),
                                        ptr(diag), ptr(b), ptr(x), int(max_iter), float(tol), ctypes.byref(it),
                                        ctypes.byref(res), int(mode))
        return SolveOutput(Status(st), x, it.value, res.value)

    # -- the solve steps as separate calls (xjpcg_ext.h; the reference's PCGImpl methods) -----------------
    def set_coo_mat(self, n, row, col, data, mode=Mode.DEFAULT) -> Status:
        """xJPCG_setCooMat: load a COO matrix (DEFAULT) or refresh its values (KEEP_NZ_LAYOUT)."""
        row = _arr(row, np.uint32)
        col = _arr(col, np.uint32)
        data = _arr(data, np.float64)
        return Status(self._L.xJPCG_setCooMat(self._h, int(n), 0 if data is None else int(data.size), ptr(row),
                                              ptr(col), ptr(data), int(mode)))

    def set_csc_sym_mat(self, n, row_idx, col_ptr, data, mode=Mode.DEFAULT) -> Status:
        """xJPCG_setCscSymMat: load a half-stored symmetric CSC matrix or refresh its values."""
        row_idx = _arr(row_idx, np.int64)
        col_ptr = _arr(col_ptr, np.int64)
        data = _arr(data, np.float64)
        return Status(self._L.xJPCG_setCscSymMat(self._h, int(n), 0 if data is None else int(data.size),
                                                 ptr(row_idx), ptr(col_ptr), ptr(data), int(mode)))

    def set_vec(self, n, b, diag) -> Status:
        """xJPCG_setVec: right-hand side and Jacobi diagonal."""
        b = _arr(b, np.float64)
        diag = _arr(diag, np.float64)
        return Status(self._L.xJPCG_setVec(self._h, int(n), ptr(b), ptr(diag)))

    def run(self, max_iter, tol):
        """xJPCG_run: solve from x0 = 0; returns (status, iterations, residual)."""
        it = ctypes.c_uint32(0)
        res = ctypes.c_double(0.0)
        st = self._L.xJPCG_run(self._h, int(max_iter), float(tol), ctypes.byref(it), ctypes.byref(res))
        return Status(st), it.value, res.value

    def results(self):
        """xJPCG_getResults: (status, iterations, residual) of the last solve."""
        it = ctypes.c_uint32(0)
        res = ctypes.c_double(0.0)
        st = self._L.xJPCG_getResults(self._h, ctypes.byref(it), ctypes.byref(res))
        return Status(st), it.value, res.value

    def get_x(self, n):
        """xJPCG_getX: (status, x) of the last solve."""
        x = np.zeros(int(n), dtype=np.float64)
        return Status(self._L.xJPCG_getX(self._h, x.ctypes.data, int(n))), x
