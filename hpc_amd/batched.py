"""Batched JPCG over include/xjpcg_batch.h: many uneven SPD systems x R right-hand sides on one GPU."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from ._lib import XjpcgError, lib, ptr


def _check(st):
    if st != 0:
        raise XjpcgError(st, lib().xjpcg_batch_last_error().decode(errors="replace"))


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@dataclass
class BatchInfo:
    launched: int                  # batched iterations launched
    iterations: torch.Tensor       # int32 [S, R] iterations per pair
    residual: torch.Tensor         # fp64 [S, R] ||r|| / ||b|| per pair


def assemble(mats, split_fp32: bool = False):
    """Block-diagonal CSR of the systems (host numpy).

    mats: objects with rowptr / col / val / diag (hpc_amd.workloads.CsrMatrix).
    With split_fp32 the off-diagonal values become fp32 and the diagonal is
    kept apart in fp64.  Returns dict(rowptr, col, val, dsplit|None, diag, sys_row, exact).
    """
    sys_row = np.zeros(len(mats) + 1, np.int64)
    for s, A in enumerate(mats):
        sys_row[s + 1] = sys_row[s] + A.n
    N = int(sys_row[-1])
    if split_fp32:
        from .device import split_fp32 as _split
        parts = [_split(A.rowptr, A.col, A.val) for A in mats]
        exact = all(p[4] for p in parts)
        cols = [p[1].astype(np.int64) + sys_row[s] for s, p in enumerate(parts)]
        vals = [p[2] for p in parts]
        counts = np.concatenate([np.diff(p[0]) for p in parts])
        dsplit = np.concatenate([p[3] for p in parts])
    else:
        exact = True
        cols = [A.col.astype(np.int64) + sys_row[s] for s, A in enumerate(mats)]
        vals = [A.val for A in mats]
        counts = np.concatenate([np.diff(A.rowptr) for A in mats])
        dsplit = None
    rowptr = np.zeros(N + 1, np.int64)
    np.cumsum(counts, out=rowptr[1:])
    if rowptr[-1] >= 2**31 or N >= 2**31:
        raise ValueError("batch too large for int32 indices")
    return {"rowptr": rowptr.astype(np.int32), "col": np.concatenate(cols).astype(np.int32),
            "val": np.concatenate(vals), "dsplit": dsplit,
            "diag": np.concatenate([A.diag for A in mats]), "sys_row": sys_row.astype(np.int32), "exact": exact}


class Batch:
    """Device batch of S systems (block-diagonal CSR in HBM) with R right-hand sides each."""

    def __init__(self, rowptr_h: np.ndarray, col: torch.Tensor, val: torch.Tensor, sys_row_h: np.ndarray, nrhs: int,
                 dsplit: torch.Tensor | None = None, device=None):
        device = device or col.device
        self.rowptr_h = np.ascontiguousarray(rowptr_h, np.int32)
        self.sys_row_h = np.ascontiguousarray(sys_row_h, np.int32)
        self.N = int(self.sys_row_h[-1])
        self.S = self.sys_row_h.size - 1
        self.R = int(nrhs)
        if self.rowptr_h.size != self.N + 1:
            raise ValueError("rowptr does not match the system row ranges")
        self.rowptr = torch.from_numpy(self.rowptr_h).to(device)
        self.col, self.val, self.dsplit = col, val, dsplit
        self.f32 = val.dtype == torch.float32
        h = ctypes.c_void_p()
        _check(lib().xjpcg_batch_create(ctypes.byref(h), self.S, self.R, self.sys_row_h.ctypes.data,
                                        self.rowptr_h.ctypes.data, ptr(self.rowptr), ptr(col), ptr(val),
                                        1 if self.f32 else 0, ptr(dsplit), _stream()))
        self._h = h

    @classmethod
    def from_host(cls, asm: dict, nrhs: int, device, fp32: bool | None = None):
        """From assemble(): fp32 values when the assembly was split (or fp32 requested)."""
        col = torch.from_numpy(asm["col"]).to(device)
        val_np = asm["val"]
        if fp32:
            val_np = val_np.astype(np.float32)
        val = torch.from_numpy(np.ascontiguousarray(val_np)).to(device)
        ds = None if asm["dsplit"] is None else torch.from_numpy(asm["dsplit"]).to(device)
        return cls(asm["rowptr"], col, val, asm["sys_row"], nrhs, dsplit=ds, device=device)

    def close(self):
        if getattr(self, "_h", None):
            lib().xjpcg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    LAYOUT_AUTO, LAYOUT_CSR, LAYOUT_DIAGONALS = 0, 1, 2

    def info(self):
        v = [ctypes.c_int32() for _ in range(5)]
        _check(lib().xjpcg_batch_info(self._h, *[ctypes.byref(a) for a in v]))
        d = dict(zip(["rows", "nnz", "tiles", "tile_rows", "block"], [a.value for a in v]))
        d["layout"] = {1: "csr", 2: "diagonals"}[self.layout]
        return d

    def set_layout(self, layout: int) -> int:
        """SpMM storage: 0 the creation-time / autotuned choice, 1 CSR, 2 per-system diagonals; returns the one
        in use."""
        act = ctypes.c_int32()
        _check(lib().xjpcg_batch_set_layout(self._h, int(layout), ctypes.byref(act),
                                            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return act.value

    @property
    def layout(self) -> int:
        act = ctypes.c_int32()
        _check(lib().xjpcg_batch_set_layout(self._h, -1, ctypes.byref(act), None))
        return act.value

    def set_vectors(self, bits: int) -> int:
        """Work vectors r, p, Ap of the solve in fp64 (bits=64) or fp32 (bits=32; x, arithmetic and dot products
        stay fp64; for the inner batch of solve_mixed); returns the setting."""
        act = ctypes.c_int32()
        _check(lib().xjpcg_batch_set_vectors(self._h, int(bits), ctypes.byref(act)))
        return act.value

    @property
    def vector_bits(self) -> int:
        act = ctypes.c_int32()
        _check(lib().xjpcg_batch_set_vectors(self._h, 0, ctypes.byref(act)))
        return act.value

    def set_path(self, path: int, grid: int = 0):
        """Iteration loop of solve(): 0 auto / 2 graph-replayed launches, 1 one persistent cooperative launch of
        `grid` workgroups (0: the default size); both give the same bits."""
        _check(lib().xjpcg_batch_set_path(self._h, int(path), int(grid), None, None))

    @property
    def last_path(self) -> tuple[int, int]:
        """(loop the last solve used, its persistent grid)"""
        p, g = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().xjpcg_batch_set_path(self._h, -1, 0, ctypes.byref(p), ctypes.byref(g)))
        return p.value, g.value

    def configure(self, tile_rows: int = 0, block: int = 0, tune: bool = False):
        _check(lib().xjpcg_batch_set_config(self._h, int(tile_rows), int(block), 1 if tune else 0))

    def solve(self, diag: torch.Tensor, b: torch.Tensor, max_iter: int, tol: float,
              x: torch.Tensor | None = None) -> tuple[torch.Tensor, BatchInfo]:
        """b, x: [N, R] fp64 (row-major); diag: [N] fp64."""
        assert b.shape == (self.N, self.R) and b.dtype == torch.float64 and b.is_contiguous()
        if x is None:
            x = torch.empty_like(b)
        its = torch.zeros((self.S, self.R), dtype=torch.int32, device=b.device)
        res = torch.zeros((self.S, self.R), dtype=torch.float64, device=b.device)
        n = ctypes.c_int64()
        _check(lib().xjpcg_batch_solve(self._h, ptr(diag), ptr(b), ptr(x), int(max_iter), float(tol), ptr(its),
                                       ptr(res), ctypes.byref(n), _stream()))
        return x, BatchInfo(n.value, its, res)


def solve_mixed(outer: Batch, inner: Batch, diag: torch.Tensor, b: torch.Tensor, tol: float, inner_tol: float,
                inner_max_iter: int = 100000, max_outer: int = 20, x: torch.Tensor | None = None):
    """fp64 iterative refinement around the (fp32-valued) inner batch; returns x, BatchInfo, outer steps."""
    if x is None:
        x = torch.empty_like(b)
    its = torch.zeros((outer.S, outer.R), dtype=torch.int32, device=b.device)
    res = torch.zeros((outer.S, outer.R), dtype=torch.float64, device=b.device)
    steps, launched = ctypes.c_uint32(), ctypes.c_int64()
    _check(lib().xjpcg_batch_solve_mixed(outer.handle, inner.handle, ptr(diag), ptr(b), ptr(x), int(max_outer),
                                         float(tol), float(inner_tol), int(inner_max_iter), ptr(its), ptr(res),
                                         ctypes.byref(steps), ctypes.byref(launched), _stream()))
    return x, BatchInfo(launched.value, its, res), steps.value
