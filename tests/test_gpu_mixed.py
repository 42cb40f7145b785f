"""Mixed-precision refined JPCG (fp32-valued inner operator, fp64 refinement) against the fp64 oracle."""
import sys

import numpy as np
import pytest

from hpc_amd import workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _ops(cuda, A):
    import torch
    from hpc_amd import device
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    full = device.CsrOperator(t(A.rowptr), t(A.col), t(A.val), rowptr_host=A.rowptr)
    srp, scol, sv32, sdiag, exact = device.split_fp32(A.rowptr, A.col, A.val)
    sd = t(sdiag)
    inner = device.CsrOperator(t(srp), t(scol), t(sv32), dsplit=sd, rowptr_host=srp)
    return full, inner, exact, t


def test_mixed_graph_laplacian_exact_fp32(cuda):
    """Graph Laplacian + 1e-2 I with fp32 weights: the split fp32 operator is A exactly, so one
    refinement step (inner tolerance 0.5 tol, since res = 1 at x0 = 0) reaches the tolerance; its
    iteration count is the fp64 oracle's at tolerance 0.5 tol, up to rounding of the split diagonal."""
    from hpc_amd import device
    A = workloads.graph_laplacian(40000, seed=2)
    b = workloads.uniform(2, 4, A.n, -1, 1)
    full, inner, exact, t = _ops(cuda, A)
    assert exact
    tol = 1e-10
    x, info = device.solve_mixed(full, inner, t(A.diag), t(b), tol)
    ref = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, b, 20000, 0.5 * tol)
    xh = x.cpu().numpy()
    err = np.linalg.norm(xh - ref["x"]) / np.linalg.norm(ref["x"])
    true = np.linalg.norm(b - A.to_scipy() @ xh) / np.linalg.norm(b)
    _say(f"graph n={A.n}: outer {info.outer} inner {info.inner_total} (oracle {ref['executed']}) "
         f"true res {true:.2e} err vs fp64 oracle {err:.2e}")
    assert info.outer == 1
    assert abs(info.inner_total - ref["executed"]) <= 3
    assert true < tol and info.residual < tol
    assert err < 1e-9


def test_mixed_inexact_fp32_values_refines(cuda):
    """exp(GRF) conductivities are not representable in fp32: the inner operator only approximates A and
    the fp64 refinement must close the gap over several steps."""
    from hpc_amd import device
    A, b, _ = workloads.c5_diffusion(m=128, length=8.0, seed=4)
    full, inner, exact, t = _ops(cuda, A)
    assert not exact
    tol = 1e-10
    x, info = device.solve_mixed(full, inner, t(A.diag), t(b), tol, inner_tol=1e-5, inner_max_iter=20000)
    ref = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, b, 50000, tol)
    import scipy.sparse.linalg as spla
    xd = spla.spsolve(A.to_scipy().tocsc(), b)
    xh = x.cpu().numpy()
    true = np.linalg.norm(b - A.to_scipy() @ xh) / np.linalg.norm(b)
    err = np.linalg.norm(xh - xd) / np.linalg.norm(xd)
    err_ref = np.linalg.norm(ref["x"] - xd) / np.linalg.norm(xd)
    _say(f"diffusion 128^2: outer {info.outer} inner {info.inner_total} (fp64 oracle {ref['executed']}) "
         f"true res {true:.2e} forward err {err:.2e} (fp64 oracle {err_ref:.2e})")
    assert info.outer >= 2
    assert true < tol and info.residual < tol
    # same residual bound => forward error of the same order as the fp64 solve's
    assert err < max(20.0 * err_ref, 1e-9)


def test_mixed_zero_rhs(cuda):
    import torch
    from hpc_amd import device
    A = workloads.laplace2d(20)
    full, inner, exact, t = _ops(cuda, A)
    x, info = device.solve_mixed(full, inner, t(A.diag), torch.zeros(A.n, dtype=torch.float64, device=cuda), 1e-10)
    assert info.outer == 0 and info.residual == 0.0 and not x.cpu().numpy().any()


def test_compact_columns_match_int32(cuda):
    """The int16 column-delta stream and the int32 stream give bit-identical SpMV results and solves."""
    import torch
    s = workloads.c2_stencil27(36)
    full, _, _, t = _ops(cuda, s.A)
    assert full.compact
    x = t(workloads.uniform(7, 1, s.A.n, -1, 1))
    y16 = full.spmv(x).clone()
    assert full.set_compact(False) is False
    y32 = full.spmv(x).clone()
    assert bool((y16 == y32).all())
    d, b = t(s.A.diag), t(s.b)
    x32, i32 = full.solve(d, b, s.max_iter, s.tol)
    x32 = x32.clone()
    assert full.set_compact(True) is True
    x16, i16 = full.solve(d, b, s.max_iter, s.tol)
    assert i16.executed == i32.executed and bool((x16 == x32).all())


def test_graph_long_rows_spmv(cuda):
    """Rows longer than a tile (the whole-workgroup path) next to short rows, with the int16 and
    the int32 column streams."""
    import torch
    from hpc_amd import device
    rng = np.random.default_rng(5)
    n = 6000
    deg = np.where(rng.random(n) < 0.01, rng.integers(2100, 9000, n), rng.integers(1, 30, n))
    rows = np.repeat(np.arange(n), deg)
    cols = rng.integers(0, n, rows.size)
    vals = rng.standard_normal(rows.size)
    rp, col, val, _ = oracle.coo_to_csr(n, rows, cols, vals)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    op = device.CsrOperator(t(rp), t(col), t(val), rowptr_host=rp)
    xv = workloads.uniform(5, 3, n, -1, 1)
    ref = oracle.spmv(rp, col, val, xv)
    short = np.diff(rp) <= 2044
    assert short.sum() < n
    for compact in (True, False):
        assert op.set_compact(compact) is compact
        y = op.spmv(t(xv)).cpu().numpy()
        assert np.array_equal(y[short], ref[short])
        assert np.allclose(y, ref, rtol=1e-12, atol=1e-10)
