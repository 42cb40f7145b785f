"""The numpy statement of the GPU kernels' row-binned summation order (tests/sum_order.py) -- CPU only.

It must reduce to the oracle's left-to-right sums on short rows, stay within the recursive-summation bound of
them on long rows, and combine lanes in the xor-butterfly order (a hand-worked case where that order and the
left-to-right order round differently)."""
import numpy as np

from hpc_amd import workloads
from oracle import oracle

import sum_order


def test_butterfly_order_hand_worked():
    # lanes 1 and 3 hold 2^-53 each, lane 0 holds 1: the butterfly adds 2^-53 + 2^-53 first (o=2, lane 1),
    # then 1 + 2^-52 (o=1, lane 0), exact; left to right rounds 1 + 2^-53 back to 1 twice
    v = np.zeros(64)
    v[0], v[1], v[3] = 1.0, 2.0 ** -53, 2.0 ** -53
    assert sum_order.butterfly(v) == 1.0 + 2.0 ** -52
    acc = 0.0
    for t in v:
        acc += t
    assert acc == 1.0
    # integers: every order is exact
    w = np.arange(64, dtype=np.float64)
    assert sum_order.butterfly(w) == 2016.0


def test_strided_partials_and_block_total():
    q = np.arange(1, 601, dtype=np.float64)
    p = sum_order.strided_partials(q, 64)
    assert p[0] == sum(q[0::64]) and p[63] == sum(q[63::64])
    assert sum_order.wave_total(q) == q.sum() == sum_order.block_total(q)


def _rows(lens, seed):
    rng = np.random.default_rng(seed)
    n = len(lens)
    rows = np.repeat(np.arange(n), lens)
    cols = np.concatenate([np.sort(rng.choice(n, d, replace=False)) for d in lens])
    vals = rng.standard_normal(rows.size)
    rp, col, val, _ = oracle.coo_to_csr(n, rows, cols, vals)
    return rp, col, val


def test_tile_order_short_rows_are_left_to_right_long_rows_within_bound():
    lens = [0, 1, 5, 63, 64, 65, 127, 128, 129, 500, 2044, 2045, 4096] * 3 + [7] * 5000
    rp, col, val = _rows(lens, 3)
    n = rp.size - 1
    x = workloads.uniform(1, 2, n, -1, 1)
    ref = oracle.spmv(rp, col, val, x)
    y = sum_order.tile_spmv(rp, col, val, x)
    L = np.diff(rp)
    assert np.array_equal(y[L <= 64], ref[L <= 64])
    assert np.all(np.abs(y - ref) <= 2 * sum_order.sum_bound(rp, col, val, x))
    assert not np.array_equal(y[L > 64], ref[L > 64])   # the bins do round differently somewhere
    d = workloads.uniform(1, 3, n, 1, 2)
    yd = sum_order.tile_spmv(rp, col, val, x, dsplit=d)
    refd = ref + 0.0
    assert np.all(np.abs(yd - (d * x + refd)) <= 2 * sum_order.sum_bound(rp, col, val, x, dsplit=d))
