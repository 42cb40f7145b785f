"""Device-level solver objects over include/xjpcg_dev.h, on torch tensors in HBM.

torch provides the device memory and the stream; every computation runs in
the HIP kernels of libxjpcg.so.  Calls launch on torch's current stream, so
torch.cuda.Event timing on that stream brackets the library's kernels.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from ._lib import check_dev, lib, ptr


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev_check(t, dtype, name, n=None):
    if not (t.is_cuda and t.dtype == dtype and t.is_contiguous()):
        raise ValueError(f"{name} must be a contiguous {dtype} CUDA tensor")
    if n is not None and t.numel() != n:
        raise ValueError(f"{name} must have {n} entries, has {t.numel()}")


@dataclass
class SolveInfo:
    iterations: int       # reported count, min(executed, maxIter)
    executed: int         # iterations run
    residual: float       # ||r_recursive|| / ||b||


@dataclass
class MixedInfo:
    outer: int            # refinement steps
    inner_total: int      # inner JPCG iterations over all steps
    residual: float       # true ||b - A x|| / ||b|| in fp64


class CsrOperator:
    """CSR operator in HBM (int32 rowptr/col, fp64 or fp32 values, optional split fp64 diagonal).

    The tensors are referenced, not copied; the operator keeps them alive.
    """

    def __init__(self, rowptr: torch.Tensor, col: torch.Tensor, val: torch.Tensor, dsplit: torch.Tensor | None = None,
                 rowptr_host=None):
        n = rowptr.numel() - 1
        _dev_check(rowptr, torch.int32, "rowptr")
        _dev_check(col, torch.int32, "col")
        if val.dtype not in (torch.float64, torch.float32):
            raise ValueError("val must be float64 or float32")
        _dev_check(val, val.dtype, "val", col.numel())
        if dsplit is not None:
            _dev_check(dsplit, torch.float64, "dsplit", n)
        self.n = n
        self.nnz = col.numel()
        self.rowptr, self.col, self.val, self.dsplit = rowptr, col, val, dsplit
        self.f32 = val.dtype == torch.float32
        h = ctypes.c_void_p()
        hp = None
        if rowptr_host is not None:
            import numpy as np
            self._rp_host = np.ascontiguousarray(rowptr_host, dtype=np.int32)
            hp = self._rp_host.ctypes.data
        check_dev(lib().xjpcg_op_create(ctypes.byref(h), n, self.nnz, ptr(rowptr), ptr(col), ptr(val),
                                        1 if self.f32 else 0, ptr(dsplit), hp, _stream()))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().xjpcg_op_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def info(self):
        n, nnz, nt, g = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        check_dev(lib().xjpcg_op_info(self._h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(nt),
                                      ctypes.byref(g)))
        return {"n": n.value, "nnz": nnz.value, "ntiles": nt.value, "grid": g.value}

    def spmv_bytes(self, layout: int = 0) -> int:
        """Bytes one SpMV launch on `layout` (0: the layout in use) moves: its matrix streams with padding, its
        per-row tables, x once and y once (-1 when that layout is not built)."""
        nb = ctypes.c_int64()
        check_dev(lib().xjpcg_op_spmv_bytes(self._h, int(layout), ctypes.byref(nb)))
        return nb.value

    def set_values(self, val: torch.Tensor):
        """Point the operator at new values (same dtype and length); pattern and tiles are kept."""
        if val.dtype != self.val.dtype:
            raise ValueError("values must keep their dtype")
        _dev_check(val, val.dtype, "val", self.nnz)
        check_dev(lib().xjpcg_op_set_values(self._h, ptr(val)))
        self.val = val

    def set_compact(self, on: bool) -> bool:
        """Use the int16 column-delta stream when it exists; returns whether it is in use."""
        act = ctypes.c_int()
        check_dev(lib().xjpcg_op_set_compact(self._h, 1 if on else 0, ctypes.byref(act)))
        return bool(act.value)

    @property
    def compact(self) -> bool:
        act = ctypes.c_int()
        check_dev(lib().xjpcg_op_set_compact(self._h, -1, ctypes.byref(act)))
        return bool(act.value)

    LAYOUT_AUTO, LAYOUT_TILES, LAYOUT_PANELS, LAYOUT_PATTERNS, LAYOUT_DIAGONALS = 0, 1, 2, 3, 4

    def set_layout(self, layout: int) -> int:
        """SpMV layout: 0 the creation-time choice, 1 row tiles, 2 panel-swept copy, 3 row tiles on column-offset
        patterns, 4 diagonals; returns the layout in use."""
        act = ctypes.c_int()
        check_dev(lib().xjpcg_op_set_layout(self._h, int(layout), _stream(), ctypes.byref(act)))
        return act.value

    @property
    def layout(self) -> int:
        act = ctypes.c_int()
        check_dev(lib().xjpcg_op_set_layout(self._h, -1, None, ctypes.byref(act)))
        return act.value

    @property
    def solve_layout(self) -> int:
        """Layout the solves run on (set_layout(L > 0) sets both; 0 restores both creation-time choices)."""
        act = ctypes.c_int()
        check_dev(lib().xjpcg_op_solve_layout(self._h, ctypes.byref(act)))
        return act.value

    def set_vectors(self, bits: int):
        """Work vectors of the solves: 64, or 32 to store r, p, A p and the Jacobi factors in fp32 (x, the
        arithmetic and the dot products stay fp64).  32 needs the diagonal or the panel solve layout."""
        check_dev(lib().xjpcg_op_set_vectors(self._h, int(bits), None))

    @property
    def vector_bits(self) -> int:
        bits = ctypes.c_int()
        check_dev(lib().xjpcg_op_set_vectors(self._h, 0, ctypes.byref(bits)))
        return bits.value

    PATH_AUTO, PATH_PERSISTENT, PATH_MULTI = 0, 1, 2

    def set_path(self, path: int):
        """Solver loop: 0 auto, 1 persistent cooperative kernel, 2 three launches per iteration."""
        check_dev(lib().xjpcg_op_set_path(self._h, int(path), None))

    @property
    def last_path(self) -> int:
        last = ctypes.c_int()
        check_dev(lib().xjpcg_op_set_path(self._h, -1, ctypes.byref(last)))
        return last.value

    def spmv(self, x: torch.Tensor, y: torch.Tensor | None = None) -> torch.Tensor:
        _dev_check(x, torch.float64, "x", self.n)
        if y is None:
            y = torch.empty_like(x)
        check_dev(lib().xjpcg_spmv(self._h, ptr(x), ptr(y), _stream()))
        return y

    def residual(self, b: torch.Tensor, x: torch.Tensor, r: torch.Tensor | None = None):
        _dev_check(b, torch.float64, "b", self.n)
        _dev_check(x, torch.float64, "x", self.n)
        if r is None:
            r = torch.empty_like(b)
        rr = ctypes.c_double()
        check_dev(lib().xjpcg_residual(self._h, ptr(b), ptr(x), ptr(r), ctypes.byref(rr), _stream()))
        return r, rr.value

    def solve(self, diag: torch.Tensor, b: torch.Tensor, max_iter: int, tol: float, x: torch.Tensor | None = None,
              records: torch.Tensor | None = None) -> tuple[torch.Tensor, SolveInfo]:
        """Jacobi PCG from x0 = 0 (reference semantics); records: uint8 tensor of (max_iter+2)*64 bytes."""
        _dev_check(diag, torch.float64, "diag", self.n)
        _dev_check(b, torch.float64, "b", self.n)
        if x is None:
            x = torch.empty_like(b)
        _dev_check(x, torch.float64, "x", self.n)
        if records is not None:
            _dev_check(records, torch.uint8, "records", (int(max_iter) + 2) * 64)
        it, res, ex = ctypes.c_uint32(), ctypes.c_double(), ctypes.c_int64()
        check_dev(lib().xjpcg_solve(self._h, ptr(diag), ptr(b), ptr(x), int(max_iter), float(tol), ctypes.byref(it),
                                    ctypes.byref(res), ctypes.byref(ex), ptr(records), _stream()))
        return x, SolveInfo(it.value, ex.value, res.value)

    def sddmm(self, lam: torch.Tensor, x: torch.Tensor, g: torch.Tensor | None = None) -> torch.Tensor:
        _dev_check(lam, torch.float64, "lam", self.n)
        _dev_check(x, torch.float64, "x", self.n)
        if g is None:
            g = torch.empty(self.nnz, dtype=torch.float64, device=x.device)
        check_dev(lib().xjpcg_sddmm(self._h, ptr(lam), ptr(x), ptr(g), _stream()))
        return g


def solve_mixed(outer: CsrOperator, inner: CsrOperator, diag: torch.Tensor, b: torch.Tensor, tol: float,
                inner_tol: float = 0.0, inner_max_iter: int = 20000, max_outer: int = 30,
                x: torch.Tensor | None = None, raise_on_fail: bool = True) -> tuple[torch.Tensor, MixedInfo]:
    """fp32-valued inner JPCG inside an fp64 iterative-refinement loop."""
    if not inner.f32:
        raise ValueError("inner operator must hold fp32 values")
    _dev_check(diag, torch.float64, "diag", outer.n)
    _dev_check(b, torch.float64, "b", outer.n)
    if x is None:
        x = torch.empty_like(b)
    no, it, res = ctypes.c_uint32(), ctypes.c_int64(), ctypes.c_double()
    st = lib().xjpcg_solve_mixed(outer.handle, inner.handle, ptr(diag), ptr(b), ptr(x), int(max_outer), float(tol),
                                 float(inner_tol), int(inner_max_iter), ctypes.byref(no), ctypes.byref(it),
                                 ctypes.byref(res), _stream())
    if st != 0 and (raise_on_fail or st != 4):
        check_dev(st)
    return x, MixedInfo(no.value, it.value, res.value)


STATIC_MAX_ROWS = 262144  # XJPCG_STATIC_MAX_ROWS


def csr_solve_static(rowptr: torch.Tensor, col: torch.Tensor, val: torch.Tensor, diag: torch.Tensor,
                     b: torch.Tensor, max_iter: int, tol: float,
                     x: torch.Tensor | None = None) -> tuple[torch.Tensor, SolveInfo]:
    """Allocation-free JPCG (xjpcg_csr_solve_static) on a device CSR with fp64 values and n <= STATIC_MAX_ROWS:
    the library's work vectors live in its static per-GPU workspace; x (allocated here when not given) is the
    only output."""
    n = rowptr.numel() - 1
    _dev_check(rowptr, torch.int32, "rowptr")
    _dev_check(col, torch.int32, "col")
    _dev_check(val, torch.float64, "val", col.numel())
    _dev_check(diag, torch.float64, "diag", n)
    _dev_check(b, torch.float64, "b", n)
    if x is None:
        x = torch.empty_like(b)
    _dev_check(x, torch.float64, "x", n)
    it, res, ex = ctypes.c_uint32(), ctypes.c_double(), ctypes.c_int64()
    check_dev(lib().xjpcg_csr_solve_static(n, col.numel(), ptr(rowptr), ptr(col), ptr(val), ptr(diag), ptr(b), ptr(x),
                                           int(max_iter), float(tol), ctypes.byref(it), ctypes.byref(res),
                                           ctypes.byref(ex), _stream()))
    return x, SolveInfo(it.value, ex.value, res.value)


def split_fp32(rowptr_h, col_h, val_h):
    """Host helper: fp32 off-diagonal CSR + fp64 diagonal from an fp64 CSR (numpy arrays).

    Returns (rowptr, col, val32, diag64, exact) where exact tells whether every
    off-diagonal value is representable in fp32 (then the split operator is the
    fp64 operator exactly).
    """
    import numpy as np
    n = rowptr_h.size - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowptr_h))
    is_d = col_h == rows
    diag = np.zeros(n, dtype=np.float64)
    np.add.at(diag, rows[is_d], val_h[is_d])
    keep = ~is_d
    cnt = np.bincount(rows[keep], minlength=n)
    rp = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(cnt, out=rp[1:])
    v32 = val_h[keep].astype(np.float32)
    exact = bool(np.all(v32.astype(np.float64) == val_h[keep]))
    return rp, np.ascontiguousarray(col_h[keep], dtype=np.int32), v32, diag, exact


def panel_plan(rowptr_h, col_h):
    """Host-only panel layout of a CSR pattern (numpy arrays), as an operator would build it.

    Returns dict(block_rows, nb, P, seg[nb*P+1], off[nnz] uint16, lrow[nnz] uint16, perm[nnz] int64, ordered).
    """
    import numpy as np
    rowptr_h = np.ascontiguousarray(rowptr_h, dtype=np.int32)
    col_h = np.ascontiguousarray(col_h, dtype=np.int32)
    n = rowptr_h.size - 1
    br, nb, P = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    check_dev(lib().xjpcg_panel_dims(n, ctypes.byref(br), ctypes.byref(nb), ctypes.byref(P)))
    nnz = int(rowptr_h[-1])
    seg = np.zeros(nb.value * P.value + 1, np.int32)
    off = np.zeros(nnz, np.uint16)
    lrow = np.zeros(nnz, np.uint16)
    perm = np.zeros(nnz, np.int64)
    ordered = ctypes.c_int()
    check_dev(lib().xjpcg_panel_plan(n, rowptr_h.ctypes.data, col_h.ctypes.data, seg.ctypes.data, off.ctypes.data,
                                     lrow.ctypes.data, perm.ctypes.data, ctypes.byref(ordered)))
    return {"block_rows": br.value, "nb": nb.value, "P": P.value, "seg": seg, "off": off, "lrow": lrow,
            "perm": perm, "ordered": bool(ordered.value)}


def tile_plan(rowptr_h):
    """Host-only row tiles of a CSR pattern: first row of each tile (ntiles+1 entries)."""
    import numpy as np
    rowptr_h = np.ascontiguousarray(rowptr_h, dtype=np.int32)
    n = rowptr_h.size - 1
    tiles = np.zeros(n + 1, np.int32)
    nt = ctypes.c_int32()
    check_dev(lib().xjpcg_tile_plan(n, rowptr_h.ctypes.data, tiles.ctypes.data, ctypes.byref(nt)))
    return tiles[:nt.value + 1].copy()


def pattern_plan(rowptr_h, col_h):
    """Host-only column-offset patterns for the row tiles: (tiles, mask[n] uint32, offs[ntiles, 32]) or None."""
    import numpy as np
    rowptr_h = np.ascontiguousarray(rowptr_h, dtype=np.int32)
    col_h = np.ascontiguousarray(col_h, dtype=np.int32)
    n = rowptr_h.size - 1
    tiles = tile_plan(rowptr_h)
    mask = np.zeros(n, np.uint32)
    offs = np.zeros((tiles.size - 1, 32), np.int32)
    ok = ctypes.c_int()
    check_dev(lib().xjpcg_pattern_plan(n, rowptr_h.ctypes.data, col_h.ctypes.data, mask.ctypes.data,
                                       offs.ctypes.data, ctypes.byref(ok)))
    return (tiles, mask, offs) if ok.value else None


def stream_copy(src: torch.Tensor, dst: torch.Tensor):
    """dst = src (fp64, same length) with the library's 16-byte non-temporal stream kernel."""
    _dev_check(src, torch.float64, "src")
    _dev_check(dst, torch.float64, "dst", src.numel())
    check_dev(lib().xjpcg_stream_copy(ptr(src), ptr(dst), src.numel(), _stream()))


def csr_from_coo(n: int, row, col, base: int = 0, device=None):
    """Device CSR from host COO index arrays (numpy uint32-compatible): (rowptr, col, perm) torch tensors.

    Rows sorted, columns ascending within a row, duplicates in input order; perm[k] is the index of entry k's
    value in the caller's value array (values on the device: val[perm])."""
    import numpy as np
    row = np.ascontiguousarray(row, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    nnz = row.size
    device = device or torch.device("cuda:0")
    rp = torch.empty(n + 1, dtype=torch.int32, device=device)
    c = torch.empty(max(nnz, 1), dtype=torch.int32, device=device)
    pm = torch.empty(max(nnz, 1), dtype=torch.int64, device=device)
    check_dev(lib().xjpcg_csr_from_coo(n, nnz, row.ctypes.data, col.ctypes.data, int(base), ptr(rp), ptr(c), ptr(pm),
                                       _stream()))
    return rp, c[:nnz], pm[:nnz]


def csr_from_cscsym(n: int, row_idx, col_ptr, base: int = 0, device=None):
    """Device CSR from a host half-stored symmetric CSC (int64 arrays): (rowptr, col, perm) torch tensors."""
    import numpy as np
    row_idx = np.ascontiguousarray(row_idx, dtype=np.int64)
    col_ptr = np.ascontiguousarray(col_ptr, dtype=np.int64)
    nh = row_idx.size
    full = max(2 * nh - n, 0)
    device = device or torch.device("cuda:0")
    rp = torch.empty(n + 1, dtype=torch.int32, device=device)
    c = torch.empty(max(full, 1), dtype=torch.int32, device=device)
    pm = torch.empty(max(full, 1), dtype=torch.int64, device=device)
    m = ctypes.c_int64()
    check_dev(lib().xjpcg_csr_from_cscsym(n, nh, row_idx.ctypes.data, col_ptr.ctypes.data, int(base), ptr(rp), ptr(c),
                                          ptr(pm), ctypes.byref(m), _stream()))
    return rp, c[:m.value], pm[:m.value]
