"""Signature files of the reference's SpMV engine, over include/xjpcg_sig.h.

The reference prepares a matrix for its FPGA SpMV as a directory of
"signature" files (L2/sparse/src/sw/c++/gen_signature.cpp:58-81): info.dat,
rbParam.dat, parParam.dat and nnzVal_0..15.dat.  `Signature` reads such a
directory (checking every size and count), writes one from a COO or a
half-stored symmetric CSC (the same bytes as the reference's partitioner,
SpmPar<double>::partitionCooMat / partitionCscSymMat, gen_signature.hpp:41-56),
rewrites values for a kept pattern (updateMat, gen_signature.hpp:61-69), and
decodes to COO or straight to a device CSR for the GPU operators.

    sig = Signature.load("matrices/nd24k")
    rowptr, col, val = sig.to_csr()            # torch tensors on cuda:0
    op = hpc_amd.device.CsrOperator(rowptr, col, val)
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import XjpcgError, lib

N_CHANNELS = 16
INFO_FIELDS = ("m", "n", "nnz", "m_pad", "n_pad", "nnz_pad")


def _check(st):
    if st != 0:
        raise XjpcgError(st, lib().xjpcg_sig_last_error().decode(errors="replace"))


class Signature:
    """One partitioned matrix (owns the library object)."""

    def __init__(self, handle: ctypes.c_void_p):
        self._h = handle

    # -- construction --------------------------------------------------------
    @classmethod
    def from_coo(cls, m: int, n: int, row, col, val, base: int = 0) -> "Signature":
        row = np.ascontiguousarray(row, dtype=np.uint32)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        if not (row.size == col.size == val.size):
            raise ValueError("row, col and val must have the same length")
        h = ctypes.c_void_p()
        _check(lib().xjpcg_sig_from_coo(ctypes.byref(h), m, n, row.size, row.ctypes.data, col.ctypes.data,
                                        val.ctypes.data, int(base)))
        return cls(h)

    @classmethod
    def from_cscsym(cls, n: int, row_idx, col_ptr, val, base: int = 0) -> "Signature":
        row_idx = np.ascontiguousarray(row_idx, dtype=np.int64)
        col_ptr = np.ascontiguousarray(col_ptr, dtype=np.int64)
        val = np.ascontiguousarray(val, dtype=np.float64)
        if val.size != row_idx.size:
            raise ValueError("row_idx and val must have the same length")
        h = ctypes.c_void_p()
        _check(lib().xjpcg_sig_from_cscsym(ctypes.byref(h), n, row_idx.size, row_idx.ctypes.data,
                                           col_ptr.ctypes.data, val.ctypes.data, int(base)))
        return cls(h)

    @classmethod
    def load(cls, directory: str) -> "Signature":
        h = ctypes.c_void_p()
        _check(lib().xjpcg_sig_load(ctypes.byref(h), os.fsencode(directory)))
        return cls(h)

    # -- lifetime ------------------------------------------------------------
    def close(self):
        if self._h is not None and self._h.value:
            lib().xjpcg_sig_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _handle(self):
        if self._h is None:
            raise ValueError("signature is closed")
        return self._h

    # -- operations ----------------------------------------------------------
    def store(self, directory: str):
        """Write the files into an existing directory."""
        _check(lib().xjpcg_sig_store(self._handle(), os.fsencode(directory)))

    def update(self, val):
        """New values for the same pattern (length of the partitioned value array)."""
        val = np.ascontiguousarray(val, dtype=np.float64)
        _check(lib().xjpcg_sig_update(self._handle(), val.ctypes.data, val.size))

    def info(self) -> dict:
        """info.dat fields plus decoded entry count, row blocks and partitions."""
        v = (ctypes.c_uint32 * 6)()
        entries = ctypes.c_uint64()
        rb = ctypes.c_uint32()
        par = ctypes.c_uint32()
        _check(lib().xjpcg_sig_info(self._handle(), v, ctypes.byref(entries), ctypes.byref(rb), ctypes.byref(par)))
        d = dict(zip(INFO_FIELDS, (int(x) for x in v)))
        d.update(entries=entries.value, row_blocks=rb.value, partitions=par.value)
        return d

    def buffer(self, which) -> bytes:
        """Raw bytes: 'rbParam', 'parParam' or a channel number 0..15 (nnzVal_<c>)."""
        idx = {"rbParam": 0, "parParam": 1}.get(which)
        if idx is None:
            idx = 2 + int(which)
        p = ctypes.c_void_p()
        nbytes = ctypes.c_uint64()
        _check(lib().xjpcg_sig_buffer(self._handle(), idx, ctypes.byref(p), ctypes.byref(nbytes)))
        return ctypes.string_at(p.value, nbytes.value) if nbytes.value else b""

    def to_coo(self):
        """(row, col, val) numpy arrays, 0-based, in stream order; zero values are not returned."""
        cnt = self.info()["entries"]
        row = np.empty(max(cnt, 1), np.uint32)
        col = np.empty(max(cnt, 1), np.uint32)
        val = np.empty(max(cnt, 1), np.float64)
        _check(lib().xjpcg_sig_to_coo(self._handle(), row.ctypes.data, col.ctypes.data, val.ctypes.data, cnt))
        return row[:cnt], col[:cnt], val[:cnt]

    def to_csr(self, device=None):
        """Device CSR (rowptr int32, col int32, val fp64 torch tensors) of a square signature, decoded on the GPU."""
        import torch
        v = (ctypes.c_uint32 * 6)()
        _check(lib().xjpcg_sig_info(self._handle(), v, None, None, None))
        m, n, nnz = int(v[0]), int(v[1]), int(v[2])
        if m != n:
            raise ValueError("to_csr needs a square matrix")
        device = device or torch.device("cuda:0")
        cap = max(nnz, 1)
        rp = torch.empty(n + 1, dtype=torch.int32, device=device)
        c = torch.empty(cap, dtype=torch.int32, device=device)
        val = torch.empty(cap, dtype=torch.float64, device=device)
        cnt = ctypes.c_uint64()
        stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
        _check(lib().xjpcg_sig_to_csr(self._handle(), rp.data_ptr(), c.data_ptr(), val.data_ptr(), cap,
                                      ctypes.byref(cnt), stream))
        k = cnt.value
        return rp, c[:k], val[:k]


def file_names(directory: str):
    """Paths of the signature files in the order the reference's SpMV test reads them (spmv/main.cpp:86-93)."""
    return ([os.path.join(directory, f"nnzVal_{c}.dat") for c in range(N_CHANNELS)]
            + [os.path.join(directory, "parParam.dat"), os.path.join(directory, "rbParam.dat"),
               os.path.join(directory, "info.dat")])
