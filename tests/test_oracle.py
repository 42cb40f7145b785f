"""Pins of the CPU oracle (oracle/jpcg_oracle.c) -- CPU only.

Each restated function is checked against something that does not restate
it: the reference's own host code compiled in oracle/_ref (bit-exact), dense
linear algebra, closed forms, and hand-worked small cases.
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def ref():
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("reference tree not available: oracle/_ref not built")
    return R


def csr_from_dense(D):
    n = D.shape[0]
    rp, col, val = [0], [], []
    for i in range(n):
        nz = np.nonzero(D[i])[0]
        col += nz.tolist()
        val += D[i, nz].tolist()
        rp.append(len(col))
    return np.array(rp, np.int32), np.array(col, np.int32), np.array(val, np.float64)


# --------------------------------------------------------------------------- ingest

@pytest.mark.parametrize("base", [0, 1])
def test_coo_to_csr_matches_reference_loadcoo(base):
    R = ref()
    rng = np.random.default_rng(1)
    n, nnz = 50, 400
    row = rng.integers(0, n, nnz).astype(np.uint32) + base
    col = rng.integers(0, n, nnz).astype(np.uint32) + base
    val = rng.standard_normal(nnz)
    r_ref = np.zeros(nnz, np.uint32)
    c_ref = np.zeros(nnz, np.uint32)
    assert R.ref_load_coo(n, nnz, row.ctypes.data, col.ctypes.data, base, r_ref.ctypes.data, c_ref.ctypes.data) == 0
    rp, c, v, perm = oracle.coo_to_csr(n, row, col, val, base)
    # every CSR slot k is input entry perm[k]; the reference's index lists after the base shift
    rows_csr = np.repeat(np.arange(n), np.diff(rp))
    assert np.array_equal(rows_csr, r_ref[perm].astype(np.int64))
    assert np.array_equal(c, c_ref[perm].astype(np.int32))
    assert np.array_equal(v, val[perm])
    # stable within a row
    for i in range(n):
        assert np.all(np.diff(perm[rp[i]:rp[i + 1]]) > 0)
    # same dense matrix, duplicates summed
    D = np.zeros((n, n))
    np.add.at(D, (r_ref.astype(np.int64), c_ref.astype(np.int64)), val)
    D2 = np.zeros((n, n))
    np.add.at(D2, (rows_csr, c.astype(np.int64)), v)
    assert np.allclose(D, D2, rtol=0, atol=1e-15)


@pytest.mark.parametrize("base", [0, 1])
@pytest.mark.parametrize("upper", [False, True])
def test_cscsym_expansion_matches_reference(base, upper):
    R = ref()
    rng = np.random.default_rng(2)
    n = 40
    M = rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.15)
    M = M + M.T + np.diag(rng.uniform(5, 6, n))
    H = np.triu(M) if upper else np.tril(M)
    import scipy.sparse as sp
    C = sp.csc_matrix(H)
    C.sort_indices()
    rowidx = C.indices.astype(np.int64) + base
    colptr = C.indptr.astype(np.int64) + base
    data = C.data.copy()
    full = 2 * data.size - n
    r_ref = np.zeros(full, np.uint32)
    c_ref = np.zeros(full, np.uint32)
    v_ref = np.zeros(full, np.float64)
    assert R.ref_cscsym_to_coo(n, full, rowidx.ctypes.data, colptr.ctypes.data, data.ctypes.data, base,
                               r_ref.ctypes.data, c_ref.ctypes.data, v_ref.ctypes.data) == 0
    r, c, v = oracle.cscsym_to_coo(n, rowidx, colptr, data, base)
    assert np.array_equal(r, r_ref) and np.array_equal(c, c_ref) and np.array_equal(v, v_ref)
    D = np.zeros((n, n))
    np.add.at(D, (r.astype(np.int64), c.astype(np.int64)), v)
    assert np.array_equal(D, M)


def test_cscsym_wrong_count_is_rejected():
    # a column without its diagonal breaks nnz_full = 2 nnz - n (the reference raises)
    rowidx = np.array([1, 1], np.int64)
    colptr = np.array([0, 1, 2], np.int64)
    with pytest.raises(ValueError):
        oracle.cscsym_to_coo(2, rowidx, colptr, np.array([1.0, 2.0]))


# --------------------------------------------------------------------------- SpMV and dot

@pytest.mark.parametrize("case", [0, 1])
def test_spmv_against_dense_fixtures(case):
    z = np.load(os.path.join(GOLDEN, "spmv_cases.npz"))
    n = int(z[f"n{case}"])
    rp, col, val, _ = oracle.coo_to_csr(n, z[f"row{case}"], z[f"col{case}"], z[f"val{case}"])
    y = oracle.spmv(rp, col, val, z[f"x{case}"])
    assert np.allclose(y, z[f"y{case}"], rtol=1e-13, atol=1e-13)
    assert y[1] == 0.0 if case == 0 else True  # empty row


def test_spmv_threads_agree():
    from hpc_amd import workloads
    A = workloads.stencil27(20)
    x = workloads.uniform(1, 2, A.n, -1, 1)
    assert np.array_equal(oracle.spmv(A.rowptr, A.col, A.val, x), oracle.spmv(A.rowptr, A.col, A.val, x, threads=4))


def test_spmv_closed_form_laplacian_row_sums():
    from hpc_amd import workloads
    m = 17
    A = workloads.laplace2d(m)
    y = oracle.spmv(A.rowptr, A.col, A.val, np.ones(A.n)).reshape(m, m)
    missing = np.zeros((m, m))
    missing[0, :] += 1
    missing[-1, :] += 1
    missing[:, 0] += 1
    missing[:, -1] += 1
    assert np.array_equal(y, missing)


def test_dot_fixed_order():
    # integers: exact, so any order gives the same value -> pins that every term is counted once
    rng = np.random.default_rng(3)
    for n in (1, 3, 4, 5, 31, 32, 33, 100, 257):
        a = rng.integers(-50, 50, n).astype(np.float64)
        b = rng.integers(-50, 50, n).astype(np.float64)
        assert oracle.dot_fixed(a, b) == float(np.dot(a, b))
    # the grouping itself: words of 4 summed (p0+p1)+(p2+p3), groups of 8 words as a tree, groups left to right
    a = np.array([1e16, 1.0, -1e16, 1.0] + [0.0] * 28 + [1.0, 1e-16] + [0.0] * 30)
    b = np.ones_like(a)
    w0 = (1e16 + 1.0) + (-1e16 + 1.0)
    g1 = 1.0 + 1e-16
    assert oracle.dot_fixed(a, b) == w0 + g1


# --------------------------------------------------------------------------- set-up and records

def test_cg_init_matches_reference_bitexact():
    R = ref()
    rng = np.random.default_rng(4)
    for n in (1, 5, 8, 1001):
        b = rng.standard_normal(n)
        d = rng.uniform(0.5, 3.0, n)
        dot, rz, da = ctypes.c_double(), ctypes.c_double(), ctypes.c_uint32()
        da_n = (n + 3) // 4 * 4
        J = np.zeros(da_n)
        z = np.zeros(da_n)
        p = np.zeros(da_n)
        R.ref_cg_init(ctypes.c_uint32(n), b.ctypes.data, d.ctypes.data, ctypes.byref(dot), ctypes.byref(rz),
                      ctypes.byref(da), J.ctypes.data, z.ctypes.data, p.ctypes.data)
        o_dot, o_rz, o_da = oracle.cg_init(b, d)
        assert (o_dot, o_rz, o_da) == (dot.value, rz.value, da.value)
        assert np.array_equal(J[n:], np.ones(da_n - n)) and np.array_equal(z[:n], (1.0 / d) * b)


def test_record_codec_matches_reference_bytes():
    R = ref()
    R.ref_record_encode.argtypes = [ctypes.c_uint32] * 4 + [ctypes.c_double] * 5 + [ctypes.c_uint64, ctypes.c_void_p]
    out = np.zeros(64, np.uint8)
    rec = dict(id=3, iter=17, vec_size=4096, max_iter=5001, tols=1.5e-20, res=2.5e-9, rz=-3.25, alpha=0.125,
               beta=7.0e-3, clock=0x0123456789ABCDEF)
    R.ref_record_encode(3, 17, 4096, 5001, 1.5e-20, 2.5e-9, -3.25, 0.125, 7.0e-3, 0x0123456789ABCDEF,
                        out.ctypes.data)
    assert oracle.record_encode(rec) == out.tobytes()
    # the oracle's decoder and the reference's read the same fields back
    assert oracle.record_decode(out.tobytes()) == rec
    u, f, clk = np.zeros(4, np.uint32), np.zeros(5), np.zeros(1, np.uint64)
    R.ref_record_decode(out.ctypes.data, u.ctypes.data, f.ctypes.data, clk.ctypes.data)
    assert [int(v) for v in u] == [3, 17, 4096, 5001]
    assert f.tolist() == [1.5e-20, 2.5e-9, -3.25, 0.125, 7.0e-3] and int(clk[0]) == 0x0123456789ABCDEF
    # record 0 as the reference host programs it (maxIter+1, Tols = dot tol^2)
    R.ref_record0.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                              ctypes.c_void_p]
    r0 = np.zeros(64, np.uint8)
    R.ref_record0(100, 64, 2.0, 1e-3, 0.5, r0.ctypes.data)
    dec = oracle.record_decode(r0.tobytes())
    assert dec == {k: v.item() for k, v in zip(oracle.RECORD_DTYPE.names, r0.view(oracle.RECORD_DTYPE)[0])}
    assert dec["max_iter"] == 101 and dec["vec_size"] == 64 and dec["res"] == 2.0 and dec["rz"] == 0.5
    assert dec["tols"] == 2.0 * 1e-3 * 1e-3
    with pytest.raises(ValueError):
        oracle.record_decode(b"\0" * 63)


# --------------------------------------------------------------------------- JPCG

def test_jpcg_dense_cholesky_fixtures():
    z = np.load(os.path.join(GOLDEN, "dense_spd.npz"))
    for t in range(3):
        D = z[f"A{t}"]
        rp, col, val = csr_from_dense(D)
        out = oracle.jpcg(rp, col, val, np.diag(D).copy(), z[f"b{t}"], 5000, 1e-14)
        x = z[f"x{t}"]
        assert np.linalg.norm(out["x"] - x) / np.linalg.norm(x) < 1e-9
        assert out["res"] < 1e-14


def test_jpcg_laplace_fixture():
    z = np.load(os.path.join(GOLDEN, "laplace2d_12.npz"))
    n = z["b"].size
    out = oracle.jpcg(z["indptr"], z["indices"], z["data"], np.full(n, 4.0), z["b"], 1000, 1e-13)
    assert np.linalg.norm(out["x"] - z["x"]) / np.linalg.norm(z["x"]) < 1e-11


def test_jpcg_diagonal_matrix_one_iteration():
    n = 37
    d = np.linspace(1.0, 9.0, n)
    b = np.sin(np.arange(n))
    out = oracle.jpcg(np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), d, d, b, 50, 1e-12)
    assert out["executed"] == 1
    assert np.allclose(out["x"], b / d, rtol=1e-15)


@pytest.mark.parametrize("k", [2, 3, 5])
def test_jpcg_k_distinct_eigenvalues(k):
    # D^-1 A with k distinct eigenvalues => CG on the preconditioned system ends in <= k steps
    rng = np.random.default_rng(10 + k)
    n = 30
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.repeat(np.linspace(1.0, 4.0, k), int(np.ceil(n / k)))[:n]
    Bm = (q * ev) @ q.T  # = D^-1/2 A D^-1/2
    dvec = rng.uniform(1.0, 3.0, n)
    Ds = np.sqrt(dvec)
    A = Ds[:, None] * Bm * Ds[None, :]
    A = 0.5 * (A + A.T)
    b = rng.standard_normal(n)
    rp, col, val = csr_from_dense(A)
    out = oracle.jpcg(rp, col, val, dvec, b, 100, 1e-10)
    assert out["executed"] <= k + 1  # exact arithmetic gives k; allow one for rounding
    assert np.linalg.norm(A @ out["x"] - b) / np.linalg.norm(b) < 1e-9


def test_jpcg_count_convention_2x2():
    # Worked by hand: A = [[4,1],[1,3]], b = [1,2], J = diag(A).  CG ends in 2 iterations exactly.
    A = np.array([[4.0, 1.0], [1.0, 3.0]])
    rp, col, val = csr_from_dense(A)
    d = np.diag(A).copy()
    b = np.array([1.0, 2.0])
    out = oracle.jpcg(rp, col, val, d, b, 10, 1e-10, records=True)
    assert out["executed"] == 2 and out["iter"] == 2
    assert np.allclose(out["x"], np.linalg.solve(A, b), rtol=1e-14)
    # maxIter = 1: the loop still runs maxIter+1 = 2 iterations (count capped at 1) -> converged residual
    out1 = oracle.jpcg(rp, col, val, d, b, 1, 1e-10, records=True)
    assert out1["executed"] == 2 and out1["iter"] == 1 and out1["res"] < 1e-10
    rec = out1["records"]
    assert rec[0]["max_iter"] == 2 and rec[1]["iter"] == 0 and rec[2]["iter"] == 1
    # first iteration by hand: z0 = J b = [1/4, 2/3], p0 = z0, rz0 = 1/4 + 4/3
    z0 = b / d
    Ap = A @ z0
    alpha = (b @ z0) / (z0 @ Ap)
    assert np.isclose(rec[1]["alpha"], alpha, rtol=1e-15)
    r1 = b - alpha * Ap
    assert np.isclose(rec[1]["res"], r1 @ r1, rtol=1e-13)
    assert np.isclose(rec[1]["beta"], (r1 @ (r1 / d)) / (b @ z0), rtol=1e-13)


def test_jpcg_zero_rhs_reference_semantics():
    # the reference divides 0/0 here (NaN residual, reported as success); the oracle restates it
    A = np.array([[2.0, 0.0], [0.0, 3.0]])
    rp, col, val = csr_from_dense(A)
    out = oracle.jpcg(rp, col, val, np.diag(A).copy(), np.zeros(2), 5, 1e-10)
    assert out["executed"] == 6 and np.isnan(out["res"])


def test_jpcg_threads_variant_close():
    from hpc_amd import workloads
    s = workloads.c1_laplace2d(48)
    a = oracle.jpcg(s.A.rowptr, s.A.col, s.A.val, s.A.diag, s.b, s.max_iter, s.tol)
    b = oracle.jpcg(s.A.rowptr, s.A.col, s.A.val, s.A.diag, s.b, s.max_iter, s.tol, threads=4)
    assert abs(a["executed"] - b["executed"]) <= 1
    assert np.linalg.norm(a["x"] - b["x"]) / np.linalg.norm(a["x"]) < 1e-10


def test_oracle_mistakes_are_caught():
    """A planted mistake in the restatement (wrong beta: using the new r.r instead of r.z) must fail the
    dense fixture check -- done on the CPU by re-running the algorithm in numpy with the mistake."""
    z = np.load(os.path.join(GOLDEN, "dense_spd.npz"))
    A, b, x_ref = z["A1"], z["b1"], z["x1"]
    d = np.diag(A)
    x = np.zeros_like(b)
    r = b.copy()
    zz = r / d
    p = zz.copy()
    rz = r @ zz
    for _ in range(200):
        Ap = A @ p
        alpha = rz / (p @ Ap)
        x += alpha * p
        r -= alpha * Ap
        zz = r / d
        rzn = r @ r  # planted mistake
        p = zz + (rzn / rz) * p
        rz = rzn
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) > 1e-6
