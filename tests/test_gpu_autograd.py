"""Differentiable solve: SDDMM gradient and adjoint JPCG against dense fp64 gradients and finite differences."""
import sys

import numpy as np
import pytest

from hpc_amd import workloads

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _setup(cuda, A):
    import torch
    from hpc_amd import autograd
    pat = autograd.SparsePattern(A.rowptr, A.col, cuda)
    vals = torch.from_numpy(A.val.copy()).to(cuda)
    return pat, vals


def test_sddmm_bitexact(cuda):
    import torch
    A = workloads.graph_laplacian(4000, seed=6)
    pat, vals = _setup(cuda, A)
    op = pat.operator(vals)
    lam = workloads.uniform(6, 1, A.n, -1, 1)
    x = workloads.uniform(6, 2, A.n, -1, 1)
    g = op.sddmm(torch.from_numpy(lam).to(cuda), torch.from_numpy(x).to(cuda)).cpu().numpy()
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    assert np.array_equal(g, -lam[rows] * x[A.col])


def test_gradient_matches_dense_fp64(cuda):
    """Small n: analytic gradient from dense fp64 solves, and per-entry (non-symmetric) central FD with
    dense LU solves, against the GPU adjoint + SDDMM."""
    import torch
    from hpc_amd import autograd
    A, b, xt = workloads.c5_diffusion(m=12, length=3.0, seed=4)
    pat, vals = _setup(cuda, A)
    vals.requires_grad_(True)
    bt = torch.from_numpy(b).to(cuda).requires_grad_(True)
    x = autograd.jpcg_solve(pat, vals, bt, tol=1e-15, max_iter=5000)
    loss = 0.5 * ((x - torch.from_numpy(xt).to(cuda)) ** 2).sum()
    loss.backward()
    gv = vals.grad.cpu().numpy()
    gb = bt.grad.cpu().numpy()
    D = A.to_scipy().toarray()
    xd = np.linalg.solve(D, b)
    lam = np.linalg.solve(D.T, xd - xt)
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    g_dense = -lam[rows] * xd[A.col]
    assert np.linalg.norm(gv - g_dense) / np.linalg.norm(g_dense) < 1e-10
    assert np.linalg.norm(gb - lam) / np.linalg.norm(lam) < 1e-10
    # non-symmetric single-entry central FD (dense LU solves), h = 1e-6
    rng = np.random.default_rng(0)
    h = 1e-6
    for k in rng.choice(A.nnz, 16, replace=False):
        i, j = rows[k], A.col[k]
        Dp, Dm = D.copy(), D.copy()
        Dp[i, j] += h
        Dm[i, j] -= h
        lp = 0.5 * np.sum((np.linalg.solve(Dp, b) - xt) ** 2)
        lm = 0.5 * np.sum((np.linalg.solve(Dm, b) - xt) ** 2)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gv[k]) <= 1e-6 * max(abs(gv[k]), 1e-3 * np.abs(gv).max()), (k, fd, gv[k])


@pytest.mark.timeout(600)
def test_c5_finite_differences_64_entries(cuda):
    """2D variable-coefficient diffusion 512^2 (exp(GRF) conductivities): gradients through the adjoint JPCG
    against central finite differences (h = 1e-6) on 64 random stored entries.  A symmetric pair (i,j),(j,i)
    is perturbed together so the perturbed operator stays SPD; its FD derivative is g_ij + g_ji.  The
    perturbed solves are warm-started as corrections x* + delta with (A +- hE) delta = b - (A +- hE) x*."""
    import torch
    from hpc_amd import autograd
    A, b, xt = workloads.c5_diffusion()
    pat, vals = _setup(cuda, A)
    tol, maxit = 1e-14, 200000
    vals.requires_grad_(True)
    bt = torch.from_numpy(b).to(cuda).requires_grad_(True)
    xt_d = torch.from_numpy(xt).to(cuda)
    x = autograd.jpcg_solve(pat, vals, bt, tol=tol, max_iter=maxit)
    loss = 0.5 * ((x - xt_d) ** 2).sum()
    loss.backward()
    _say(f"C5 forward iterations {x.grad_fn.info.executed if hasattr(x.grad_fn, 'info') else '?'}")
    gv = vals.grad.detach().clone()
    xs = x.detach().clone()
    tidx = pat.transpose_index()
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    rng = np.random.default_rng(4)
    picks = rng.choice(A.nnz, 64, replace=False)
    h = 1e-6
    base = vals.detach().clone()
    errs = []
    for k in picks:
        t = int(tidx[k])
        deltas = []
        for sgn in (1.0, -1.0):
            pv = base.clone()
            pv[k] += sgn * h
            if t != k:
                pv[t] += sgn * h
            op = pat.operator(pv)
            r, _ = op.residual(bt.detach(), xs)
            d, info = op.solve(pv[pat.diag_index].contiguous(), r, maxit, 1e-13)
            deltas.append(d.clone())
        # L+ - L- = sum (d+ - d-) (x* - t + (d+ + d-)/2): formed from the corrections, elementwise,
        # so two large loss values are never subtracted
        dp, dm = deltas
        fd = float(((dp - dm) * (xs - xt_d + 0.5 * (dp + dm))).sum()) / (2 * h)
        an = float(gv[k] + (gv[t] if t != k else 0.0))
        errs.append(abs(fd - an) / max(abs(an), 1e-12))
        _say(f"entry {k} ({rows[k]},{A.col[k]}): fd {fd:.12e} analytic {an:.12e} rel {errs[-1]:.2e}")
    pat.operator(base)
    errs = np.array(errs)
    _say(f"FD relative errors: median {np.median(errs):.2e} max {errs.max():.2e}")
    assert np.median(errs) < 1e-7
    assert errs.max() < 1e-6


def test_unconverged_forward_raises(cuda):
    import torch
    from hpc_amd import autograd
    A, b, _ = workloads.c5_diffusion(m=32, length=4.0, seed=4)
    pat, vals = _setup(cuda, A)
    vals.requires_grad_(True)
    bt = torch.from_numpy(b).to(cuda).requires_grad_(True)
    with pytest.raises(autograd.SolveNotConverged) as ei:
        autograd.jpcg_solve(pat, vals, bt, tol=1e-12, max_iter=5)
    assert ei.value.which == "forward" and ei.value.status == 4
    _say(str(ei.value))


def test_unconverged_adjoint_raises(cuda):
    """Forward right-hand side = an eigenvector of the 5-point Laplacian (constant diagonal, so Jacobi CG
    converges in one iteration); the adjoint right-hand side x - x_target is not, and 5 iterations cannot
    reach the tolerance: backward must raise instead of returning a gradient."""
    import torch
    from hpc_amd import autograd
    m = 24
    A = workloads.laplace2d(m)
    s = np.sin(np.pi * np.arange(1, m + 1) / (m + 1))
    b = np.outer(s, s).ravel()
    pat, vals = _setup(cuda, A)
    vals.requires_grad_(True)
    bt = torch.from_numpy(b).to(cuda).requires_grad_(True)
    x = autograd.jpcg_solve(pat, vals, bt, tol=1e-10, max_iter=5)
    xt = torch.from_numpy(workloads.uniform(4, 2, A.n, -1.0, 1.0)).to(cuda)
    loss = 0.5 * ((x - xt) ** 2).sum()
    with pytest.raises(autograd.SolveNotConverged) as ei:
        loss.backward()
    assert ei.value.which == "adjoint"
    _say(str(ei.value))
