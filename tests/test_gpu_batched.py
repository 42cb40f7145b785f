"""Batched JPCG (many uneven systems x R right-hand sides) against the CPU oracle, per pair."""
import sys

import numpy as np
import pytest

from hpc_amd import workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def small_batch(count=6, nrhs=4, nmin=300, nmax=3000, seed=3):
    mats, rhs = workloads.c4_systems(count, nmin, nmax, nrhs, seed)
    return mats, rhs


def _dev(cuda, a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(cuda)


def test_batched_fp64_matches_oracle_per_pair(cuda):
    from hpc_amd import batched
    mats, rhs = small_batch()
    R = rhs[0].shape[1]
    asm = batched.assemble(mats)
    B = batched.Batch.from_host(asm, R, cuda)
    _say(f"batch S={B.S} N={B.N} R={R} info={B.info()}")
    b = _dev(cuda, np.concatenate(rhs))
    diag = _dev(cuda, asm["diag"])
    tol, maxit = 1e-8, 5000
    x, info = B.solve(diag, b, maxit, tol)
    xh = x.cpu().numpy()
    its = info.iterations.cpu().numpy()
    res = info.residual.cpu().numpy()
    for s, A in enumerate(mats):
        r0, r1 = asm["sys_row"][s], asm["sys_row"][s + 1]
        for j in range(R):
            ref = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, rhs[s][:, j].copy(), maxit, tol)
            assert abs(int(its[s, j]) - ref["executed"]) <= 1, (s, j, its[s, j], ref["executed"])
            e = np.linalg.norm(xh[r0:r1, j] - ref["x"]) / np.linalg.norm(ref["x"])
            assert e <= 1e-10, (s, j, e)
            assert res[s, j] < tol
    assert info.launched >= its.max()


def test_batched_zero_rhs_and_masks(cuda):
    from hpc_amd import batched
    mats, rhs = small_batch(count=3, nrhs=2)
    rhs[1][:, 0] = 0.0
    asm = batched.assemble(mats)
    B = batched.Batch.from_host(asm, 2, cuda)
    x, info = B.solve(_dev(cuda, asm["diag"]), _dev(cuda, np.concatenate(rhs)), 4000, 1e-8)
    its = info.iterations.cpu().numpy()
    assert its[1, 0] == 0
    xh = x.cpu().numpy()
    assert not xh[asm["sys_row"][1]:asm["sys_row"][2], 0].any()
    assert its[1, 1] > 0


def test_batched_iteration_cap(cuda):
    from hpc_amd import batched
    mats, rhs = small_batch(count=2, nrhs=1)
    asm = batched.assemble(mats)
    B = batched.Batch.from_host(asm, 1, cuda)
    x, info = B.solve(_dev(cuda, asm["diag"]), _dev(cuda, np.concatenate(rhs)), 7, 1e-12)
    assert np.all(info.iterations.cpu().numpy() == 8)  # maxIter+1 rounds, as the reference control loop


def test_batched_autotune_keeps_results(cuda):
    from hpc_amd import batched
    mats, rhs = small_batch(count=4, nrhs=8)
    asm = batched.assemble(mats)
    B = batched.Batch.from_host(asm, 8, cuda)
    b, d = _dev(cuda, np.concatenate(rhs)), _dev(cuda, asm["diag"])
    x1, i1 = B.solve(d, b, 5000, 1e-8)
    x1 = x1.clone()
    B.configure(tune=True)
    _say(f"autotuned: {B.info()}")
    x2, i2 = B.solve(d, b, 5000, 1e-8)
    assert np.array_equal(i1.iterations.cpu().numpy(), i2.iterations.cpu().numpy()) or \
        np.abs(i1.iterations.cpu().numpy() - i2.iterations.cpu().numpy()).max() <= 1
    assert float(((x1 - x2).norm() / x1.norm()).item()) < 1e-10
    x3, i3 = B.solve(d, b, 5000, 1e-8)
    assert bool((x2 == x3).all())  # deterministic for a fixed configuration


def test_batched_mixed_refined(cuda):
    from hpc_amd import batched
    mats, rhs = small_batch(count=5, nrhs=4, seed=8)
    asm64 = batched.assemble(mats)
    asm32 = batched.assemble(mats, split_fp32=True)
    assert asm32["exact"]
    outer = batched.Batch.from_host(asm32, 4, cuda)            # fp32 values are exact: the split operator is A
    inner = batched.Batch.from_host(asm32, 4, cuda)
    inner.configure(tune=True)
    b, d = _dev(cuda, np.concatenate(rhs)), _dev(cuda, asm64["diag"])
    tol = 1e-8
    x, info, steps = batched.solve_mixed(outer, inner, d, b, tol, inner_tol=1e-4)
    _say(f"mixed batched: outer steps {steps}, launched {info.launched}")
    assert steps >= 2  # inner 1e-4 cannot reach 1e-8 in one step
    res = info.residual.cpu().numpy()
    assert res.max() < tol
    xh = x.cpu().numpy()
    for s, A in enumerate(mats):
        r0, r1 = asm64["sys_row"][s], asm64["sys_row"][s + 1]
        M = A.to_scipy()
        for j in range(4):
            bj = rhs[s][:, j]
            true = np.linalg.norm(bj - M @ xh[r0:r1, j]) / np.linalg.norm(bj)
            assert true < tol * 1.01, (s, j, true)


@pytest.mark.parametrize("split", [False, True], ids=["fp64", "fp32-split"])
def test_batched_diagonals_match_csr(cuda, split):
    """Per-system diagonals (at most 8 offsets per system): the SpMM sums each row in CSR order, so solves on
    both layouts are identical, pair for pair; a system of more than 8 offsets refuses the layout."""
    from hpc_amd import batched
    from hpc_amd._lib import XjpcgError
    mats, rhs = small_batch(count=5, nrhs=4, seed=11)
    asm = batched.assemble(mats, split_fp32=split)
    B = batched.Batch.from_host(asm, 4, cuda)
    _say(f"diagonals: creation-time layout {B.layout} info={B.info()}")
    b = _dev(cuda, np.concatenate(rhs))
    diag = _dev(cuda, asm["diag"])
    out = {}
    for lay in (1, 2):
        assert B.set_layout(lay) == lay
        x, info = B.solve(diag, b, 5000, 1e-8)
        out[lay] = (x.cpu().numpy(), info.iterations.cpu().numpy())
        if split:  # refinement: the residual SpMM runs on the outer batch's layout
            outer = batched.Batch.from_host(asm, 4, cuda)
            assert outer.set_layout(lay) == lay
            xm, _, steps = batched.solve_mixed(outer, B, diag, b, 1e-8, inner_tol=0.0)
            out[lay] += (xm.cpu().numpy(), steps)
    assert np.array_equal(out[1][1], out[2][1])
    assert np.array_equal(out[1][0], out[2][0])
    if split:
        assert out[1][3] == out[2][3] and np.array_equal(out[1][2], out[2][2])
    # an extra coupling beyond the system's stencil: 9 offsets in one system
    A0 = mats[0]
    rows = np.repeat(np.arange(A0.n), np.diff(A0.rowptr))
    cols = A0.col.astype(np.int64)
    vals = A0.val.copy()
    extra_r = np.array([0, A0.n - 1, 5, 17])
    extra_c = np.array([A0.n // 2, A0.n // 3, 5 + 37, 17 + 101])
    rows = np.concatenate([rows, extra_r, extra_c])
    cols = np.concatenate([cols, extra_c, extra_r])
    vals = np.concatenate([vals, np.full(8, -1e-3)])
    o = np.lexsort((cols, rows))
    rp = np.zeros(A0.n + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=A0.n), out=rp[1:])
    A1 = workloads.CsrMatrix(rp.astype(np.int32), cols[o].astype(np.int32), vals[o], A0.diag)
    B2 = batched.Batch.from_host(batched.assemble([A1] + mats[1:]), 4, cuda)
    assert B2.layout == 1
    with pytest.raises(XjpcgError):
        B2.set_layout(2)


def test_batched_fp32_vectors_alone(cuda):
    """fp32 work vectors (fp64 arithmetic and dots) on their own, to a tolerance fp32 storage can carry: every
    pair stops within a few iterations of the oracle's fp64 JPCG and its true residual stays near tol."""
    from hpc_amd import batched
    mats, rhs = small_batch(count=4, nrhs=4)
    R = 4
    asm = batched.assemble(mats)
    B = batched.Batch.from_host(asm, R, cuda)
    assert B.set_vectors(32) == 32 and B.vector_bits == 32
    tol, maxit = 1e-5, 5000
    b = np.concatenate(rhs)
    x, info = B.solve(_dev(cuda, asm["diag"]), _dev(cuda, b), maxit, tol)
    xh = x.cpu().numpy()
    its = info.iterations.cpu().numpy()
    for s, A in enumerate(mats):
        r0, r1 = asm["sys_row"][s], asm["sys_row"][s + 1]
        for j in range(R):
            ref = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, rhs[s][:, j].copy(), maxit, tol)
            true = np.linalg.norm(rhs[s][:, j] - oracle.spmv(A.rowptr, A.col, A.val, xh[r0:r1, j].copy())) / \
                np.linalg.norm(rhs[s][:, j])
            _say(f"fp32 vectors: system {s} rhs {j}: {its[s, j]} it (oracle {ref['executed']}), true res {true:.2e}")
            assert abs(int(its[s, j]) - ref["executed"]) <= max(3, ref["executed"] // 20)
            assert true < 2 * tol


def test_batched_mixed_fp32_vectors_matches_oracle(cuda):
    """Refinement with an inner batch on fp32 values and fp32 work vectors: every pair reaches the true
    relative residual tol in fp64, in at least two steps (each step is asked for at most 1e-7), and lands within
    cond-bounded distance of the oracle's fp64 solution; the fp64-vector batch is unchanged by the switch."""
    from hpc_amd import batched
    mats, rhs = small_batch(count=5, nrhs=8)
    R = 8
    asm64 = batched.assemble(mats)
    asm32 = batched.assemble(mats, split_fp32=True)
    outer = batched.Batch.from_host(asm32, R, cuda)
    inner = batched.Batch.from_host(asm32, R, cuda)
    inner.set_vectors(32)
    diag = _dev(cuda, asm64["diag"])
    b = np.concatenate(rhs)
    tol = 1e-8
    xm, minfo, steps = batched.solve_mixed(outer, inner, diag, _dev(cuda, b), tol, inner_tol=0.0)
    xmh = xm.cpu().numpy()
    res = minfo.residual.cpu().numpy()
    _say(f"mixed with fp32 vectors: {steps} steps, {minfo.launched} launched, worst true res {res.max():.2e}")
    assert steps >= 2 and np.all(res < tol)
    for s, A in enumerate(mats):
        r0, r1 = asm64["sys_row"][s], asm64["sys_row"][s + 1]
        for j in (0, R - 1):
            ref = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, rhs[s][:, j].copy(), 5000, tol)
            e = np.linalg.norm(xmh[r0:r1, j] - ref["x"]) / np.linalg.norm(ref["x"])
            assert e <= 1e-5, (s, j, e)
    # the same batch back on fp64 vectors gives the fp64 solve exactly
    inner.set_vectors(64)
    B64 = batched.Batch.from_host(asm32, R, cuda)
    x1, i1 = inner.solve(diag, _dev(cuda, b), 5000, tol)
    x2, i2 = B64.solve(diag, _dev(cuda, b), 5000, tol)
    assert np.array_equal(x1.cpu().numpy(), x2.cpu().numpy())
    assert np.array_equal(i1.iterations.cpu().numpy(), i2.iterations.cpu().numpy())


# ----------------------------------------------------------------------------- one persistent launch

def _both_paths(cuda, B, diag, b, maxit, tol, grid=0):
    out = {}
    for path in (2, 1):
        B.set_path(path, grid)
        x, info = B.solve(diag, b, maxit, tol)
        out[path] = (x.cpu().numpy(), info.iterations.cpu().numpy(), info.residual.cpu().numpy(), info.launched)
        assert B.last_path[0] == path
    return out


@pytest.mark.parametrize("nrhs,split,w32,layout,block,grid", [
    (4, False, False, 1, 256, 0),      # CSR, fp64
    (4, False, False, 2, 256, 8),      # diagonals, a grid of 8 workgroups sweeping many tiles each
    (16, True, False, 2, 512, 0),      # fp32 values + fp64 split diagonal, 16 right-hand sides
    (8, True, True, 1, 128, 24),       # fp32 values and fp32 work vectors
    (1, False, False, 2, 64, 0),       # one right-hand side, 64-thread workgroups
], ids=["csr-fp64", "dia-grid8", "dia-split-r16", "csr-w32-b128", "r1-b64"])
def test_batched_persistent_same_bits_as_launches(cuda, nrhs, split, w32, layout, block, grid):
    """The iteration loop as one cooperative launch gives the launch-per-step loop's bits: x, every pair's
    iteration count and recursive residual, for every value / vector type, layout, block size and grid size;
    its iteration count is the last pair's (the launches round up to their replay chunk)."""
    from hpc_amd import batched
    mats, rhs = small_batch(count=7, nrhs=nrhs, seed=5)
    asm = batched.assemble(mats, split_fp32=split)
    B = batched.Batch.from_host(asm, nrhs, cuda)
    B.configure(tile_rows=128, block=block)
    assert B.set_layout(layout) == layout
    if w32:
        B.set_vectors(32)
    diag, b = _dev(cuda, asm["diag"]), _dev(cuda, np.concatenate(rhs))
    tol = 1e-5 if w32 else 1e-8
    out = _both_paths(cuda, B, diag, b, 5000, tol, grid)
    its = out[1][1]
    _say(f"persistent: grid {B.last_path[1]}, {out[1][3]} iterations (launches: {out[2][3]}), pair iterations "
         f"{its.min()}..{its.max()}")
    assert np.array_equal(out[1][0], out[2][0])
    assert np.array_equal(out[1][1], out[2][1])
    assert np.array_equal(out[1][2], out[2][2])
    assert out[1][3] == its.max() and out[2][3] >= its.max()
    assert np.all(out[1][2] < tol)
    if grid:
        assert B.last_path[1] == grid


def test_batched_persistent_cap_zero_rhs_and_oracle(cuda):
    """Persistent loop: the maxIter+1 cap ends the launch, a zero right-hand side stays x = 0 without running,
    and every pair matches the oracle's JPCG (iterations +-1, relative error <= 1e-10)."""
    from hpc_amd import batched
    mats, rhs = small_batch(count=4, nrhs=2, seed=9)
    rhs[2][:, 1] = 0.0
    asm = batched.assemble(mats)
    B = batched.Batch.from_host(asm, 2, cuda)
    B.set_path(1)
    diag, b = _dev(cuda, asm["diag"]), _dev(cuda, np.concatenate(rhs))
    x, info = B.solve(diag, b, 7, 1e-14)
    assert info.launched == 8 and B.last_path[0] == 1
    its = info.iterations.cpu().numpy()
    assert its[2, 1] == 0 and np.all(np.delete(its.ravel(), 2 * 2 + 1) == 8)
    tol, maxit = 1e-9, 5000
    x, info = B.solve(diag, b, maxit, tol)
    xh, its = x.cpu().numpy(), info.iterations.cpu().numpy()
    assert not xh[asm["sys_row"][2]:asm["sys_row"][3], 1].any() and its[2, 1] == 0
    for s, A in enumerate(mats):
        r0, r1 = asm["sys_row"][s], asm["sys_row"][s + 1]
        for j in range(2):
            if s == 2 and j == 1:
                continue
            ref = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, rhs[s][:, j].copy(), maxit, tol)
            assert abs(int(its[s, j]) - ref["executed"]) <= 1, (s, j)
            e = np.linalg.norm(xh[r0:r1, j] - ref["x"]) / np.linalg.norm(ref["x"])
            assert e <= 1e-10, (s, j, e)


def test_batched_persistent_inner_of_refinement(cuda):
    """The inner batch of a refinement on the persistent loop: the same refined solution, bit for bit, as
    with graph-replayed launches."""
    from hpc_amd import batched
    mats, rhs = small_batch(count=5, nrhs=4, seed=8)
    asm32 = batched.assemble(mats, split_fp32=True)
    asm64 = batched.assemble(mats)
    outer = batched.Batch.from_host(asm32, 4, cuda)
    inner = batched.Batch.from_host(asm32, 4, cuda)
    inner.set_vectors(32)
    b, d = _dev(cuda, np.concatenate(rhs)), _dev(cuda, asm64["diag"])
    res = {}
    for path in (2, 1):
        inner.set_path(path)
        x, info, steps = batched.solve_mixed(outer, inner, d, b, 1e-8, inner_tol=0.0)
        res[path] = (x.cpu().numpy(), steps, info.residual.cpu().numpy())
        assert inner.last_path[0] == path
    assert res[1][1] == res[2][1] and np.array_equal(res[1][0], res[2][0])
    assert np.all(res[1][2] < 1e-8)


def test_batched_path_arguments(cuda):
    from hpc_amd import batched
    from hpc_amd._lib import XjpcgError
    mats, rhs = small_batch(count=2, nrhs=1)
    B = batched.Batch.from_host(batched.assemble(mats), 1, cuda)
    assert B.last_path == (0, 0)
    for bad in ((3, 0), (1, -1)):
        with pytest.raises(XjpcgError):
            B.set_path(*bad)


@pytest.mark.parametrize("nrhs", [2, 32])
def test_batched_persistent_more_rhs_and_oversized_grid(cuda, nrhs):
    """R = 2 and R = 32 on the persistent loop, with a grid request far above the tile count (capped to what
    the batch can use): the launch loop's bits."""
    from hpc_amd import batched
    mats, rhs = small_batch(count=3, nrhs=nrhs, nmin=500, nmax=2000, seed=13)
    asm = batched.assemble(mats)
    B = batched.Batch.from_host(asm, nrhs, cuda)
    diag, b = _dev(cuda, asm["diag"]), _dev(cuda, np.concatenate(rhs))
    out = _both_paths(cuda, B, diag, b, 5000, 1e-8, grid=100000)
    _say(f"R={nrhs}: persistent grid {B.last_path[1]} for {B.info()['tiles']} tiles")
    assert 0 < B.last_path[1] <= max(B.info()["tiles"], B.S)
    assert np.array_equal(out[1][0], out[2][0]) and np.array_equal(out[1][1], out[2][1])
    assert np.all(out[1][2] < 1e-8)


def test_batched_persistent_all_zero_rhs(cuda):
    """No pair runs: the persistent launch leaves after one masked iteration with x = 0 and no pair
    iterations."""
    import torch
    from hpc_amd import batched
    mats, rhs = small_batch(count=2, nrhs=4)
    asm = batched.assemble(mats)
    B = batched.Batch.from_host(asm, 4, cuda)
    B.set_path(1)
    b = torch.zeros((B.N, 4), dtype=torch.float64, device=cuda)
    x, info = B.solve(_dev(cuda, asm["diag"]), b, 100, 1e-8)
    assert B.last_path[0] == 1 and info.launched == 1
    assert not x.cpu().numpy().any() and not info.iterations.cpu().numpy().any()
    assert not info.residual.cpu().numpy().any()
