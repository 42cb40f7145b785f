"""Differentiable sparse SPD solve x = A(values)^-1 b through the GPU JPCG.

Forward: Jacobi PCG (x0 = 0) with A's diagonal as the preconditioner.
Backward (A symmetric positive definite, so A^T = A):
    adjoint solve   A lam = dL/dx            (the same GPU JPCG)
    dL/db           = lam
    dL/dvalues[k]   = -lam[row_k] * x[col_k]  (SDDMM kernel on the sparsity pattern)
Every stored entry is an independent parameter: the gradient of a symmetric
pair (i,j), (j,i) is split between its two entries.
A forward or adjoint solve that stops at max_iter with its recursive residual
still at or above tol raises SolveNotConverged: a gradient from an
unconverged adjoint would be silently wrong.
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import XjpcgError
from .device import CsrOperator

XJPCG_STATUS_EXECUTION_FAILED = 4


class SolveNotConverged(XjpcgError):
    """A solve inside the differentiable function ended above its tolerance (status EXECUTION_FAILED, as
    pcg.h reports a divergent solve)."""

    def __init__(self, which: str, info, tol: float, max_iter: int):
        super().__init__(XJPCG_STATUS_EXECUTION_FAILED,
                         f"{which} JPCG stopped after {info.executed} iterations (max_iter {max_iter}) at relative "
                         f"residual {info.residual:.3e}, tolerance {tol:.3e}")
        self.which, self.info = which, info


def _converged(which: str, info, tol: float, max_iter: int):
    if not (info.residual < tol):
        raise SolveNotConverged(which, info, tol, max_iter)


class SparsePattern:
    """Fixed CSR pattern on the GPU (int32 row pointers / columns) and the operator built on it."""

    def __init__(self, rowptr_h: np.ndarray, col_h: np.ndarray, device):
        self.rowptr_h = np.ascontiguousarray(rowptr_h, np.int32)
        self.col_h = np.ascontiguousarray(col_h, np.int32)
        self.n = self.rowptr_h.size - 1
        self.nnz = self.col_h.size
        if self.col_h.size and (self.col_h.min() < 0 or self.col_h.max() >= self.n):
            raise ValueError("column index out of range")
        rows = np.repeat(np.arange(self.n), np.diff(self.rowptr_h))
        dpos = np.nonzero(self.col_h == rows)[0]
        if dpos.size != self.n or np.any(rows[dpos] != np.arange(self.n)):
            raise ValueError("every row must store exactly one diagonal entry")
        self.device = device
        self.rowptr = torch.from_numpy(self.rowptr_h).to(device)
        self.col = torch.from_numpy(self.col_h).to(device)
        self.diag_index = torch.from_numpy(dpos.astype(np.int64)).to(device)
        self._op = None

    def operator(self, values: torch.Tensor) -> CsrOperator:
        if self._op is None:
            self._op = CsrOperator(self.rowptr, self.col, values, rowptr_host=self.rowptr_h)
        else:
            self._op.set_values(values)
        return self._op

    def transpose_index(self) -> np.ndarray:
        """t[k] = position of entry (col_k, row_k) (host; pattern must be structurally symmetric)."""
        rows = np.repeat(np.arange(self.n), np.diff(self.rowptr_h))
        key = rows.astype(np.int64) * self.n + self.col_h
        order = np.argsort(key, kind="stable")
        tkey = self.col_h.astype(np.int64) * self.n + rows
        pos = np.searchsorted(key[order], tkey)
        t = order[np.minimum(pos, self.nnz - 1)]
        if not np.array_equal(key[t], tkey):
            raise ValueError("pattern is not structurally symmetric")
        return t


class JPCGSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, pattern: SparsePattern, tol: float, max_iter: int):
        values = values.contiguous()
        b = b.contiguous()
        op = pattern.operator(values)
        diag = values[pattern.diag_index].contiguous()
        x, info = op.solve(diag, b, max_iter, tol)
        _converged("forward", info, tol, max_iter)
        x = x.clone()
        ctx.pattern, ctx.tol, ctx.max_iter = pattern, tol, max_iter
        ctx.info = info
        ctx.save_for_backward(values, x)
        return x

    @staticmethod
    def backward(ctx, gx):
        values, x = ctx.saved_tensors
        pattern = ctx.pattern
        op = pattern.operator(values)
        diag = values[pattern.diag_index].contiguous()
        lam, info = op.solve(diag, gx.contiguous(), ctx.max_iter, ctx.tol)
        _converged("adjoint", info, ctx.tol, ctx.max_iter)
        lam = lam.clone()
        gvals = op.sddmm(lam, x) if ctx.needs_input_grad[0] else None
        gb = lam if ctx.needs_input_grad[1] else None
        return gvals, gb, None, None, None


def jpcg_solve(pattern: SparsePattern, values: torch.Tensor, b: torch.Tensor, tol: float = 1e-12,
               max_iter: int = 100000) -> torch.Tensor:
    """x = A(values)^-1 b, differentiable in values (all stored entries) and b."""
    return JPCGSolve.apply(values, b, pattern, tol, max_iter)
