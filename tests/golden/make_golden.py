"""Writes the small golden fixtures under tests/golden/ with numpy / scipy only.

The reference commits no golden vectors (its tests download SuiteSparse
matrices and draw unseeded right-hand sides), so the fixtures pin the solve
with direct factorisations instead:

  dense_spd.npz     three dense SPD systems (n = 5, 24, 60), A = Q diag(ev) Q^T with
                    ev log-spaced in [1, 1e3]; x by Cholesky (scipy.linalg.cho_solve)
  laplace2d_12.npz  5-point Dirichlet Laplacian on 12x12 built with Kronecker
                    products; x by sparse LU (scipy.sparse.linalg.spsolve)
  spmv_cases.npz    COO matrices with duplicates, empty rows and one long row;
                    y = dense A @ x

Run: python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))


def dense_spd():
    out = {}
    for t, n in enumerate((5, 24, 60)):
        rng = np.random.default_rng(100 + t)
        q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        ev = np.logspace(0, 3, n)
        A = (q * ev) @ q.T
        A = 0.5 * (A + A.T)
        b = rng.uniform(-1, 1, n)
        x = sla.cho_solve(sla.cho_factor(A), b)
        out[f"A{t}"] = A
        out[f"b{t}"] = b
        out[f"x{t}"] = x
    np.savez(os.path.join(HERE, "dense_spd.npz"), **out)


def laplace2d_12():
    m = 12
    T = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(m, m))
    I = sp.identity(m)
    A = (sp.kron(I, T) + sp.kron(T, I)).tocsr()
    A.sort_indices()
    rng = np.random.default_rng(7)
    b = rng.uniform(-1, 1, m * m)
    x = spla.spsolve(A.tocsc(), b)
    np.savez(os.path.join(HERE, "laplace2d_12.npz"), indptr=A.indptr.astype(np.int32),
             indices=A.indices.astype(np.int32), data=A.data, b=b, x=x)


def spmv_cases():
    out = {}
    rng = np.random.default_rng(11)
    # case 0: 9x9 with duplicates and two empty rows
    n = 9
    row = np.array([0, 0, 0, 2, 2, 3, 5, 5, 5, 6, 8, 8, 3, 2], dtype=np.uint32)
    col = np.array([0, 3, 0, 1, 2, 8, 5, 4, 5, 0, 8, 7, 3, 1], dtype=np.uint32)
    val = rng.uniform(-2, 2, row.size)
    x = rng.uniform(-1, 1, n)
    D = np.zeros((n, n))
    np.add.at(D, (row.astype(np.int64), col.astype(np.int64)), val)
    out.update(n0=n, row0=row, col0=col, val0=val, x0=x, y0=D @ x)
    # case 1: 300x300 with one row of 5000 entries (repeated columns) and sparse rest
    n = 300
    r = [np.full(5000, 17)]
    c = [rng.integers(0, n, 5000)]
    r.append(rng.integers(0, n, 2000))
    c.append(rng.integers(0, n, 2000))
    row = np.concatenate(r).astype(np.uint32)
    col = np.concatenate(c).astype(np.uint32)
    val = rng.uniform(-1, 1, row.size)
    x = rng.uniform(-1, 1, n)
    D = np.zeros((n, n))
    np.add.at(D, (row.astype(np.int64), col.astype(np.int64)), val)
    out.update(n1=n, row1=row, col1=col, val1=val, x1=x, y1=D @ x)
    np.savez(os.path.join(HERE, "spmv_cases.npz"), **out)


if __name__ == "__main__":
    dense_spd()
    laplace2d_12()
    spmv_cases()
    print("fixtures written to", HERE)
