"""Writes tests/golden/signature_cases.npz: small matrices and the signature files
the reference's own partitioner writes for them.

The files come from oracle/_ref/libxjpcg_ref.so (the reference's
SpmPar<double> partitioner and Signature::store_* writers, compiled from its
sources by `make -C oracle ref`), so the fixture pins the signature encoder
and decoder where the reference tree is absent.  Cases:

  stencil  5-point matrix on a 20 x 20 grid, values N(0,1), COO in row order, base 0
  rect     9000 x 13000 random COO (600 entries, 40 repeated (row, col) pairs,
           every 9th value 0), shuffled, Fortran base 1: 3 row blocks, several
           column partitions, padding-only channels
  longrow  40 x 30000: 3000 entries, one row holding 1200 of them (rows split
           over partitions and channels, index words past 512 entries)
  cscsym   lower triangle (with diagonal) of the stencil case as a half-stored
           symmetric CSC, partitioned, then values replaced by the reference's
           value-only update (updateCscSymMat)

Each case stores its inputs and, per file, the bytes as uint8 (key
"<case>/<file>").  Run: python tests/golden/make_signature_golden.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402


def stencil(g, rng):
    idx = np.arange(g * g).reshape(g, g)
    r, c = [], []
    for i in range(g):
        for j in range(g):
            for di, dj in ((0, 0), (-1, 0), (0, -1), (0, 1), (1, 0)):
                if 0 <= i + di < g and 0 <= j + dj < g:
                    r.append(idx[i, j])
                    c.append(idx[i + di, j + dj])
    r = np.array(r, np.uint32)
    c = np.array(c, np.uint32)
    return r, c, rng.standard_normal(r.size)


def store(out, name, directory, **inputs):
    for k, v in inputs.items():
        out[f"{name}/{k}"] = np.asarray(v)
    for fn, b in oracle.read_sig_files(directory).items():
        out[f"{name}/{fn}"] = np.frombuffer(b, np.uint8)


def main():
    if oracle.ref_lib() is None:
        raise SystemExit("oracle/_ref is not built (needs the reference tree): make -C oracle ref")
    out = {}
    rng = np.random.default_rng(20261203)

    g = 20
    r, c, v = stencil(g, rng)
    with tempfile.TemporaryDirectory() as d:
        oracle.ref_sig_coo(d, g * g, g * g, r, c, v)
        store(out, "stencil", d, m=g * g, n=g * g, row=r, col=c, val=v, base=0)

    m, n, k = 9000, 13000, 600
    rr = rng.integers(0, m, k).astype(np.uint32)
    cc = rng.integers(0, n, k).astype(np.uint32)
    rr[:40] = rr[300:340]
    cc[:40] = cc[300:340]
    vv = rng.standard_normal(k)
    vv[::9] = 0.0
    p = rng.permutation(k)
    rr, cc, vv = rr[p] + 1, cc[p] + 1, vv[p]
    with tempfile.TemporaryDirectory() as d:
        oracle.ref_sig_coo(d, m, n, rr, cc, vv, base=1)
        store(out, "rect", d, m=m, n=n, row=rr, col=cc, val=vv, base=1)

    m, n, k = 40, 30000, 3000
    rr = rng.integers(0, m, k).astype(np.uint32)
    rr[:1200] = 17
    cc = rng.integers(0, n, k).astype(np.uint32)
    vv = rng.standard_normal(k)
    with tempfile.TemporaryDirectory() as d:
        oracle.ref_sig_coo(d, m, n, rr, cc, vv)
        store(out, "longrow", d, m=m, n=n, row=rr, col=cc, val=vv, base=0)

    import scipy.sparse as sp
    r, c, _ = stencil(g, rng)
    A = sp.coo_matrix((np.ones(r.size), (r, c)), shape=(g * g, g * g)).tocsc()
    L = sp.tril(A).tocsc()
    L.sort_indices()
    ri = L.indices.astype(np.int64)
    cp = L.indptr.astype(np.int64)
    v1 = rng.standard_normal(L.nnz)
    v2 = rng.standard_normal(L.nnz)
    with tempfile.TemporaryDirectory() as d:
        oracle.ref_sig_cscsym(d, g * g, ri, cp, v1, v2)
        store(out, "cscsym", d, n=g * g, row_idx=ri, col_ptr=cp, val=v1, val2=v2, base=0)

    path = os.path.join(HERE, "signature_cases.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
