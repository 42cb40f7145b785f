"""Edge cases of the SpMV in every layout that applies (row tiles, panel-swept copy, column-offset patterns,
diagonals):
a 1x1 matrix, empty rows, rows around the 2044-entry tile limit, odd entry counts (the final partial
16-byte run), an operator with no stored entries, and the solves they allow."""
import sys
import zlib

import numpy as np
import pytest

from hpc_amd import workloads
from oracle import oracle

import sum_order

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _t(cuda, a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(cuda)


def _csr(n, rows, cols, vals):
    o = np.lexsort((cols, rows))
    rows, cols, vals = rows[o], cols[o], vals[o]
    rp = np.zeros(n + 1, np.int32)
    np.cumsum(np.bincount(rows, minlength=n), out=rp[1:])
    return rp, cols.astype(np.int32), vals.astype(np.float64)


def case(name):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    if name == "n1":
        return np.array([0, 1], np.int32), np.array([0], np.int32), np.array([4.0])
    if name == "empty_rows":
        n = 3000
        deg = rng.integers(1, 20, n)
        deg[::3] = 0
        rows = np.repeat(np.arange(n), deg)
        cols = rng.choice(n, rows.size)
        return _csr(n, rows, cols, rng.standard_normal(rows.size))
    if name == "tile_limit_rows":
        n = 5000
        deg = rng.integers(1, 12, n)
        for i, d in zip((10, 500, 900, 1500, 4999), (2043, 2044, 2045, 2046, 4096)):
            deg[i] = d
        rows = np.repeat(np.arange(n), deg)
        cols = np.concatenate([np.sort(rng.choice(n, d, replace=False)) for d in deg])
        return _csr(n, rows, cols, rng.standard_normal(rows.size))
    if name == "odd_nnz":
        A = workloads.stencil(workloads.STENCIL_2D5, 31, 33, 1, [1.0, 2.0])
        assert A.nnz % 2 == 1
        return A.rowptr, A.col, A.val
    raise ValueError(name)


LAYOUTS = {1: "tiles", 2: "panels", 3: "patterns", 4: "diagonals"}


@pytest.mark.parametrize("name", ["n1", "empty_rows", "tile_limit_rows", "odd_nnz"])
@pytest.mark.parametrize("f32", [False, True], ids=["fp64", "fp32"])
def test_spmv_edges_every_layout(cuda, name, f32):
    from hpc_amd import device
    from hpc_amd._lib import XjpcgError
    rp, col, val = case(name)
    n = rp.size - 1
    if f32:
        val = val.astype(np.float32).astype(np.float64)   # fp32-exact, so the oracle sees the same matrix
    op = device.CsrOperator(_t(cuda, rp), _t(cuda, col), _t(cuda, val.astype(np.float32) if f32 else val),
                            rowptr_host=rp)
    xv = workloads.uniform(3, 1, n, -1, 1)
    ref = oracle.spmv(rp, col, val, xv)
    short = np.diff(rp) <= sum_order.WAVE_MIN
    bound = 2 * sum_order.sum_bound(rp, col, val, xv)
    ran = []
    for lay, lname in LAYOUTS.items():
        try:
            assert op.set_layout(lay) == lay
        except XjpcgError:
            continue
        y = op.spmv(_t(cuda, xv)).cpu().numpy()
        ran.append(lname)
        if lname == "tiles":
            assert np.array_equal(y[short], ref[short]), lname   # lane per row: left to right
            assert np.array_equal(y, sum_order.tile_spmv(rp, col, val, xv)), lname
        elif lname == "panels":
            assert np.array_equal(y, sum_order.panel_spmv(device.panel_plan(rp, col), n, val, xv)), lname
        else:
            assert np.array_equal(y, ref), lname        # left to right in every row
        assert np.all(np.abs(y - ref) <= bound), lname
    _say(f"{name} {'fp32' if f32 else 'fp64'}: layouts {ran}")
    assert "tiles" in ran and "panels" in ran


def test_no_stored_entries_split_diagonal(cuda):
    """An fp32 operator with no off-diagonal entries at all: y = dsplit * x, and the mixed solve of the
    diagonal system takes one step of one inner iteration."""
    import torch
    from hpc_amd import device
    n = 777
    d = workloads.uniform(5, 1, n, 1, 3)
    rp = np.zeros(n + 1, np.int32)
    empty_c = torch.zeros(4, dtype=torch.int32, device=cuda)[:0]
    empty_v = torch.zeros(4, dtype=torch.float32, device=cuda)[:0]
    inner = device.CsrOperator(_t(cuda, rp), empty_c, empty_v, dsplit=_t(cuda, d), rowptr_host=rp)
    outer = device.CsrOperator(_t(cuda, rp), empty_c, empty_v, dsplit=_t(cuda, d), rowptr_host=rp)
    xv = workloads.uniform(5, 2, n, -1, 1)
    assert np.array_equal(inner.spmv(_t(cuda, xv)).cpu().numpy(), d * xv)
    b = workloads.uniform(5, 3, n, -1, 1)
    x, info = device.solve_mixed(outer, inner, _t(cuda, d), _t(cuda, b), 1e-12)
    assert info.outer == 1 and info.inner_total == 1
    assert np.allclose(x.cpu().numpy(), b / d, rtol=1e-15, atol=0)


def test_one_by_one_solve(cuda):
    from hpc_amd import device
    rp, col, val = case("n1")
    op = device.CsrOperator(_t(cuda, rp), _t(cuda, col), _t(cuda, val), rowptr_host=rp)
    x, info = op.solve(_t(cuda, np.array([4.0])), _t(cuda, np.array([2.0])), 10, 1e-12)
    assert info.executed == 1 and x.cpu().numpy()[0] == 0.5 and info.residual == 0.0
