"""Host plan of the column-offset pattern tiles (xjpcg_pattern_plan) -- CPU only.

Decoding every row's mask against its tile's offset table must give back the row's CSR columns exactly, in
order; operators the layout cannot hold (scattered graphs, unsorted or long rows) must be refused.
"""
import numpy as np
import pytest

from hpc_amd import device, workloads


def decode(tiles, mask, offs, n):
    rows, cols = [], []
    for t in range(tiles.size - 1):
        for i in range(tiles[t], tiles[t + 1]):
            m = int(mask[i])
            b = 0
            while m:
                if m & 1:
                    rows.append(i)
                    cols.append(i + int(offs[t, b]))
                m >>= 1
                b += 1
    return np.array(rows, np.int64), np.array(cols, np.int64)


@pytest.mark.parametrize("make", [lambda: workloads.laplace2d(37), lambda: workloads.stencil27(13),
                                  lambda: workloads.stencil(workloads.STENCIL_3D7, 9, 7, 5, [1.0, 2.0, 3.0])],
                         ids=["2d5", "3d27", "3d7-aniso"])
def test_pattern_plan_decodes_to_csr(make):
    A = make()
    plan = device.pattern_plan(A.rowptr, A.col)
    assert plan is not None
    tiles, mask, offs = plan
    assert tiles[0] == 0 and tiles[-1] == A.n and np.all(np.diff(tiles) > 0)
    for t in range(tiles.size - 1):
        nb = int(np.bitwise_or.reduce(mask[tiles[t]:tiles[t + 1]])).bit_length()
        assert np.all(np.diff(offs[t, :nb]) > 0)          # ascending distinct offsets
    rows, cols = decode(tiles, mask, offs, A.n)
    assert np.array_equal(rows, np.repeat(np.arange(A.n), np.diff(A.rowptr)))
    assert np.array_equal(cols, A.col)


def test_pattern_plan_refuses_what_it_cannot_hold():
    G = workloads.graph_laplacian(5000, seed=3)
    assert device.pattern_plan(G.rowptr, G.col) is None   # far more than 32 offsets per tile
    A = workloads.laplace2d(10)
    col = A.col.copy()
    col[A.rowptr[5]], col[A.rowptr[5] + 1] = col[A.rowptr[5] + 1], col[A.rowptr[5]]
    assert device.pattern_plan(A.rowptr, col) is None     # a row out of order
    n = 100
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = 33
    rp[0] = 0
    col = np.arange(33, dtype=np.int32)
    assert device.pattern_plan(rp, col) is None           # a row of 33 entries
