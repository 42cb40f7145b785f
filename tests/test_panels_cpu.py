"""Host plan of the panel-swept SpMV layout (xjpcg_panel_plan) -- CPU only.

The plan must be a permutation of the CSR entries grouped by (row block, column panel), keep each row's
entries in CSR order, and encode every entry's row and column exactly; a numpy emulation of the kernel's
summation order over the plan must equal the oracle's left-to-right row sums bit for bit.
"""
import numpy as np
import pytest

from hpc_amd import device, workloads
from hpc_amd._lib import XjpcgError
from oracle import oracle

import sum_order


def scattered_csr(n, seed, max_deg=40):
    rng = np.random.default_rng(seed)
    deg = rng.integers(0, max_deg, n)
    deg[rng.random(n) < 0.001] = 3000          # a few rows longer than a tile
    rows = np.repeat(np.arange(n), deg)
    cols = rng.integers(0, n, rows.size)
    vals = rng.standard_normal(rows.size)
    o = np.lexsort((cols, rows))                 # ascending columns inside each row
    rp, col, val, _ = oracle.coo_to_csr(n, rows[o], cols[o], vals[o])
    assert np.all(np.diff(col)[np.repeat(np.arange(n), np.diff(rp))[1:] == np.repeat(np.arange(n), np.diff(rp))[:-1]] >= 0)
    return rp, col, val


@pytest.mark.parametrize("n", [1, 5000, 150_000, 300_000])
def test_panel_dims(n):
    pl = device.panel_plan(np.zeros(n + 1, np.int32), np.zeros(0, np.int32))
    br = pl["block_rows"]
    assert 256 <= br <= 4096 and br & (br - 1) == 0
    assert br == 256 or n >= br * 512          # blocks stay plentiful
    assert br == 4096 or n < 2 * br * 512      # and as large as that allows
    assert pl["nb"] == -(-n // br) and pl["P"] == -(-n // 65536)
    assert pl["seg"].size == pl["nb"] * pl["P"] + 1 and not pl["seg"].any()


def test_panel_plan_invariants_and_emulated_sums():
    n = 300_000
    rp, col, val = scattered_csr(n, 11)
    pl = device.panel_plan(rp, col)
    assert pl["ordered"]
    br, P, seg, perm = pl["block_rows"], pl["P"], pl["seg"], pl["perm"]
    nnz = col.size
    assert seg[0] == 0 and seg[-1] == nnz and np.all(np.diff(seg) >= 0)
    assert np.array_equal(np.sort(perm), np.arange(nnz))
    csr_row = np.repeat(np.arange(n), np.diff(rp))
    # segment of every layout entry -> (block, panel)
    q_seg = np.repeat(np.arange(seg.size - 1), np.diff(seg))
    b, p = q_seg // P, q_seg % P
    word = pl["lrow"].astype(np.int64)
    rows = b * br + (word & 0xFFF)
    cols = p * 65536 + pl["off"].astype(np.int64)
    assert np.array_equal(rows, csr_row[perm])
    assert np.array_equal(cols, col[perm])
    # row-major inside a segment, and each row's entries in CSR order across segments
    same = q_seg[1:] == q_seg[:-1]
    assert np.all(rows[1:][same] >= rows[:-1][same])
    order = np.lexsort((np.arange(nnz), rows))
    assert np.all(np.diff(perm[order])[rows[order][1:] == rows[order][:-1]] > 0)
    # run codes: inside each staged chunk of 2048 entries of a segment, the first entry of a row's run
    # carries min(run length, 15) above the 12-bit row offset, every other entry 0
    code = word >> 12
    pos = np.arange(nnz) - seg[q_seg]
    start = (pos % 2048 == 0) | np.concatenate(([True], rows[1:] != rows[:-1])) | np.concatenate(([True], same == 0))
    assert np.all(code[~start] == 0) and np.all(code[start] > 0)
    run_id = np.cumsum(start) - 1
    run_len = np.bincount(run_id)
    assert np.array_equal(code[start], np.minimum(run_len, 15))
    assert run_len.max() > 15                    # the long rows take the wavefront path
    # emulate the kernel: per row, products added in layout order starting from 0 -- left to right, the
    # oracle's order, wherever no run takes the wavefront path
    x = workloads.uniform(3, 1, n, -1, 1)
    y = np.zeros(n)
    np.add.at(y, rows, val[perm] * x[cols])
    ref = oracle.spmv(rp, col, val, x)
    assert np.array_equal(y, ref)
    # the kernel's binned order: equal to it for rows whose runs are all short, within the summation bound
    # for the others
    yb = sum_order.panel_spmv(pl, n, val, x)
    wide_rows = np.unique(rows[start][code[start] >= 15])
    short = np.ones(n, bool)
    short[wide_rows] = False
    assert wide_rows.size > 0 and np.array_equal(yb[short], ref[short])
    assert np.all(np.abs(yb - ref) <= 2 * sum_order.sum_bound(rp, col, val, x))


def test_panel_plan_flags_unordered_rows():
    n = 140_000
    rp = np.array([0, 2] + [2] * (n - 1), np.int32)
    col = np.array([100_000, 5], np.int32)       # panel 1 before panel 0 in row 0
    assert not device.panel_plan(rp, col)["ordered"]
    assert device.panel_plan(rp, np.sort(col))["ordered"]


def test_panel_plan_rejects_bad_columns():
    n = 1000
    rp = np.array([0, 1] + [1] * (n - 1), np.int32)
    with pytest.raises(XjpcgError) as e:
        device.panel_plan(rp, np.array([n], np.int32))
    assert e.value.status == 3 and "out of range" in str(e.value)  # XJPCG_STATUS_INVALID_VALUE
