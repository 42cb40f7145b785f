"""ctypes binding of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker / CPU baseline.  The product package (hpc_amd/) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libxjpcg_ref.so")

_L = None
_R = None


def build(ref: bool = False):
    targets = ["all"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        _L = ctypes.CDLL(LIB)
        v, i32, i64, u32, d = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_double
        _L.orc_coo_to_csr.argtypes = [u32, u32, v, v, v, ctypes.c_int, v, v, v, v]
        _L.orc_cscsym_to_coo.restype = i64
        _L.orc_cscsym_to_coo.argtypes = [i64, i64, v, v, v, ctypes.c_int, v, v, v]
        _L.orc_spmv.argtypes = [i32, v, v, v, v, v]
        _L.orc_spmv_omp.argtypes = [i32, v, v, v, v, v, ctypes.c_int]
        _L.orc_dot_fixed.restype = d
        _L.orc_dot_fixed.argtypes = [i64, v, v]
        _L.orc_cg_init.argtypes = [u32, v, v, ctypes.POINTER(d), ctypes.POINTER(d), ctypes.POINTER(u32)]
        _L.orc_record_encode.argtypes = [v, v]
        _L.orc_record_decode.argtypes = [v, v]
        _L.orc_jpcg.argtypes = [i32, v, v, v, v, v, u32, d, v, ctypes.POINTER(u32), ctypes.POINTER(d), v, v,
                                ctypes.c_int]
    return _L


def ref_lib():
    """oracle/_ref/libxjpcg_ref.so (reference host code), or None where it cannot be built."""
    global _R
    if _R is None:
        if not os.path.exists(REF_LIB):
            try:
                build(ref=True)
            except Exception:
                return None
        if not os.path.exists(REF_LIB):
            return None
        R = ctypes.CDLL(REF_LIB)
        v, u32, i, d = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_double
        R.ref_cg_init.argtypes = [u32, v, v, ctypes.POINTER(d), ctypes.POINTER(d), ctypes.POINTER(u32), v, v, v]
        R.ref_record0.argtypes = [u32, u32, d, d, d, v]
        R.ref_record_encode.argtypes = [u32] * 4 + [d] * 5 + [ctypes.c_uint64, v]
        R.ref_record_decode.argtypes = [v, v, v, v]
        R.ref_load_coo.argtypes = [u32, u32, v, v, i, v, v]
        R.ref_cscsym_to_coo.argtypes = [u32, u32, v, v, v, i, v, v, v]
        cs = ctypes.c_char_p
        R.ref_sig_coo.argtypes = [u32, u32, u32, v, v, v, v, i, cs, cs, i]
        R.ref_sig_cscsym.argtypes = [u32, u32, v, v, v, v, i, cs, cs, i]
        _R = R
    return _R


def _p(a):
    return None if a is None else a.ctypes.data


def coo_to_csr(n, row, col, val, base=0):
    row = np.ascontiguousarray(row, np.uint32)
    col = np.ascontiguousarray(col, np.uint32)
    val = np.ascontiguousarray(val, np.float64)
    rp = np.zeros(n + 1, np.int32)
    c = np.zeros(row.size, np.int32)
    v = np.zeros(row.size, np.float64)
    perm = np.zeros(row.size, np.int64)
    rc = lib().orc_coo_to_csr(n, row.size, _p(row), _p(col), _p(val), base, _p(rp), _p(c), _p(v), _p(perm))
    if rc != 0:
        raise ValueError("bad COO input")
    return rp, c, v, perm


def cscsym_to_coo(n, row_idx, col_ptr, data, base=0):
    row_idx = np.ascontiguousarray(row_idx, np.int64)
    col_ptr = np.ascontiguousarray(col_ptr, np.int64)
    data = np.ascontiguousarray(data, np.float64)
    full = 2 * data.size - n
    r = np.zeros(full, np.uint32)
    c = np.zeros(full, np.uint32)
    v = np.zeros(full, np.float64)
    m = lib().orc_cscsym_to_coo(n, data.size, _p(row_idx), _p(col_ptr), _p(data), base, _p(r), _p(c), _p(v))
    if m < 0:
        raise ValueError("bad CSC-sym input")
    return r, c, v


def spmv(rowptr, col, val, x, threads=1):
    n = rowptr.size - 1
    y = np.zeros(n, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    if threads > 1:
        lib().orc_spmv_omp(n, _p(rowptr), _p(col), _p(val), _p(x), _p(y), threads)
    else:
        lib().orc_spmv(n, _p(rowptr), _p(col), _p(val), _p(x), _p(y))
    return y


def dot_fixed(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return lib().orc_dot_fixed(a.size, _p(a), _p(b))


def cg_init(b, diag):
    b = np.ascontiguousarray(b, np.float64)
    diag = np.ascontiguousarray(diag, np.float64)
    dot, rz, da = ctypes.c_double(), ctypes.c_double(), ctypes.c_uint32()
    lib().orc_cg_init(b.size, _p(b), _p(diag), ctypes.byref(dot), ctypes.byref(rz), ctypes.byref(da))
    return dot.value, rz.value, da.value


RECORD_DTYPE = np.dtype([("id", "<u4"), ("iter", "<u4"), ("vec_size", "<u4"), ("max_iter", "<u4"),
                         ("tols", "<f8"), ("res", "<f8"), ("rz", "<f8"), ("alpha", "<f8"), ("beta", "<f8"),
                         ("clock", "<u8")])


def record_encode(rec: dict) -> bytes:
    a = np.zeros(1, RECORD_DTYPE)
    for k, v in rec.items():
        a[k] = v
    out = np.zeros(64, np.uint8)
    lib().orc_record_encode(a.ctypes.data, out.ctypes.data)
    return out.tobytes()


def record_decode(raw: bytes) -> dict:
    """One 64-byte record as a dict of its fields (CGSolverInstr::load and decode,
    pcg/hw/include/impl/cgInstr.hpp:94-97 and 133-148)."""
    inp = np.frombuffer(bytes(raw), np.uint8)
    if inp.size != 64:
        raise ValueError("a record is 64 bytes")
    a = np.zeros(1, RECORD_DTYPE)
    lib().orc_record_decode(inp.ctypes.data, a.ctypes.data)
    return {k: a[k][0].item() for k in RECORD_DTYPE.names}


def jpcg(rowptr, col, val, diag, b, max_iter, tol, threads=1, records=False):
    """Returns dict(x, iter, res, executed, rr_hist, records)."""
    n = rowptr.size - 1
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    diag = np.ascontiguousarray(diag, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n, np.float64)
    it, res = ctypes.c_uint32(), ctypes.c_double()
    rec = np.zeros((max_iter + 2) * 64, np.uint8) if records else None
    hist = np.zeros(max_iter + 1, np.float64)
    k = lib().orc_jpcg(n, _p(rowptr), _p(col), _p(val), _p(diag), _p(b), max_iter, tol, _p(x), ctypes.byref(it),
                       ctypes.byref(res), _p(rec), _p(hist), threads)
    if k < 0:
        raise MemoryError("oracle allocation failed")
    out = {"x": x, "iter": it.value, "res": res.value, "executed": k, "rr_hist": hist[:k]}
    if records:
        out["records"] = rec.view(RECORD_DTYPE)
    return out


# ---------------------------------------------------------------- signature files
#
# The reference prepares matrices for its FPGA SpMV as "signature" files
# (L2/sparse/src/sw/c++/gen_signature.cpp:58-81).  ref_sig_* run the
# reference's own partitioner (oracle/_ref); sig_decode restates the
# reference's own decoder in numpy.

SIG_FILES = ["info.dat", "rbParam.dat", "parParam.dat"] + [f"nnzVal_{c}.dat" for c in range(16)]


def ref_sig_coo(directory, m, n, row, col, val, val2=None, base=0):
    """Reference SpmPar<double>::partitionCooMat (+ updateMat with val2), files written by its Signature::store_*."""
    R = ref_lib()
    row = np.ascontiguousarray(row, np.uint32)
    col = np.ascontiguousarray(col, np.uint32)
    val = np.ascontiguousarray(val, np.float64)
    val2 = None if val2 is None else np.ascontiguousarray(val2, np.float64)
    err = ctypes.create_string_buffer(512)
    rc = R.ref_sig_coo(m, n, row.size, _p(row), _p(col), _p(val), _p(val2), base, os.fsencode(directory), err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode(errors="replace"))


def ref_sig_cscsym(directory, n, row_idx, col_ptr, val, val2=None, base=0):
    """Reference SpmPar<double>::partitionCscSymMat (+ updateCscSymMat with val2)."""
    R = ref_lib()
    row_idx = np.ascontiguousarray(row_idx, np.int64)
    col_ptr = np.ascontiguousarray(col_ptr, np.int64)
    val = np.ascontiguousarray(val, np.float64)
    val2 = None if val2 is None else np.ascontiguousarray(val2, np.float64)
    full = 2 * row_idx.size - n
    err = ctypes.create_string_buffer(512)
    rc = R.ref_sig_cscsym(n, full, _p(row_idx), _p(col_ptr), _p(val), _p(val2), base, os.fsencode(directory), err,
                          512)
    if rc != 0:
        raise RuntimeError(err.value.decode(errors="replace"))


def read_sig_files(directory):
    """{file name: bytes} of a signature directory."""
    out = {}
    for nm in SIG_FILES:
        with open(os.path.join(directory, nm), "rb") as f:
            out[nm] = f.read()
    return out


def sig_decode(files):
    """Entries (row, col, val) of signature files ({name: bytes}), channel by channel in stream order.

    Restates the reference decoder: L2/sparse/tests/fp64/spmv/python/signature.py:373-409
    (create_spm: per row block and channel, start row = block row + channel row
    offset; per partition, start 4-column block = partition block + channel block
    offset; values equal to 0 dropped) and matrix_params.py:300-323 (get_chPar: a
    row-index word every 512 entries, one row per 32; a column word every 64, one
    4-column block per 4 entries; a 4-double value word per 4 entries).  Small
    inputs only (a Python loop over words)."""
    W = 32
    u32 = lambda b, o: int(np.frombuffer(b, np.uint32, 1, o)[0])  # noqa: E731
    u16 = lambda b, o: int(np.frombuffer(b, np.uint16, 1, o)[0])  # noqa: E731
    rb, par = files["rbParam.dat"], files["parParam.dat"]
    n_rb = u32(rb, 4)
    rows, cols, vals = [], [], []
    lane = np.arange(4, dtype=np.int64)
    for c in range(16):
        st = files[f"nnzVal_{c}.dat"]
        off = W
        p0 = 0
        for r in range(n_rb):
            rbo = W + r * 8 * W
            row0 = u32(rb, rbo) + u16(rb, rbo + 2 * W + 2 * c)
            npars = u32(rb, rbo + 12)
            for p in range(p0, p0 + npars):
                po = W + p * 8 * W
                cnt = u32(par, po + 2 * W + 4 * c)
                blk0 = u32(par, po + 4 * W) + u16(par, po + 5 * W + 2 * c)
                for i in range(0, cnt, 4):
                    if i % 512 == 0:
                        ridx = np.frombuffer(st, np.uint16, 16, off)
                        off += W
                    if i % 64 == 0:
                        cidx = np.frombuffer(st, np.uint16, 16, off)
                        off += W
                    v = np.frombuffer(st, np.float64, 4, off)
                    off += W
                    keep = v != 0
                    rows.append(np.full(int(keep.sum()), int(ridx[(i % 512) // 32]) + row0, np.int64))
                    cols.append(((int(cidx[(i % 64) // 4]) + blk0) * 4 + lane)[keep])
                    vals.append(v[keep])
            p0 += npars
        assert off == len(st), f"channel {c}: decoded {off} of {len(st)} bytes"
    cat = lambda a, t: np.concatenate(a).astype(t) if a else np.zeros(0, t)  # noqa: E731
    return cat(rows, np.uint32), cat(cols, np.uint32), cat(vals, np.float64)
