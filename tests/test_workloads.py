"""Seeded workload generators (C++ in libxjpcg.so) against independent constructions -- CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp

from hpc_amd import workloads


def test_rng_numpy_mirror_bitexact():
    for seed, stream in [(0, 0), (1, 0), (3, 7), (2**40 + 5, 123)]:
        a = workloads.uniform(seed, stream, 5000, -1.0, 1.0)
        b = workloads.uniform_np(seed, stream, 5000, -1.0, 1.0)
        assert np.array_equal(a, b)
    u = workloads.uniform(0, 0, 200000, 0.0, 1.0)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 5e-3


def test_splitmix_known_answer():
    # SplitMix64 finaliser of the published reference stream: seed 0's state advances by the golden
    # gamma, so mix64(gamma) is its first output (0xE220A8397B1DCDAF in Vigna's splitmix64.c)
    assert int(workloads._mix64(np.uint64(0x9E3779B97F4A7C15))) == 0xE220A8397B1DCDAF


def _kron_laplace2d(nx, ny, kx, ky):
    Tx = sp.diags([-kx, 2 * kx, -kx], [-1, 0, 1], shape=(nx, nx))
    Ty = sp.diags([-ky, 2 * ky, -ky], [-1, 0, 1], shape=(ny, ny))
    return (sp.kron(sp.identity(ny), Tx) + sp.kron(Ty, sp.identity(nx))).tocsr()


@pytest.mark.parametrize("nx,ny,kx,ky", [(256, 256, 1.0, 1.0), (7, 13, 0.5, 2.25), (1, 5, 1.0, 1.0)])
def test_stencil_2d5_matches_kron(nx, ny, kx, ky):
    A = workloads.stencil(workloads.STENCIL_2D5, nx, ny, 1, [kx, ky])
    K = _kron_laplace2d(nx, ny, kx, ky)
    K.eliminate_zeros()
    K.sort_indices()
    if nx == 1:  # a single column: x-couplings vanish, diag stays 2(kx+ky)
        assert np.allclose(A.diag, 2 * (kx + ky))
        return
    assert np.array_equal(A.rowptr, K.indptr) and np.array_equal(A.col, K.indices)
    assert np.array_equal(A.val, K.data)
    assert A.nnz == 5 * nx * ny - 2 * nx - 2 * ny


def test_c1_size_and_rhs():
    s = workloads.c1_laplace2d(256)
    assert (s.A.n, s.A.nnz) == (65536, 326656)
    assert np.array_equal(s.b, s.A.to_scipy() @ s.x_true) or np.allclose(s.b, s.A.to_scipy() @ s.x_true,
                                                                          rtol=0, atol=1e-14)


def test_stencil27_structure():
    m = 9
    A = workloads.stencil27(m)
    T = sp.diags([1, 1, 1], [-1, 0, 1], shape=(m, m))
    K = sp.kron(sp.kron(T, T), T).tocsr()
    K.sort_indices()
    assert np.array_equal(A.rowptr, K.indptr) and np.array_equal(A.col, K.indices)
    diag = A.col == np.repeat(np.arange(A.n), np.diff(A.rowptr))
    assert np.all(A.val[diag] == 26.0) and np.all(A.val[~diag] == -1.0)
    assert workloads.lib().xjpcg_stencil_nnz(2, 160, 160, 160) == 478 ** 3


def test_stencil_3d7():
    A = workloads.stencil(workloads.STENCIL_3D7, 4, 5, 6, [1.0, 2.0, 3.0])
    M = A.to_scipy()
    assert (M != M.T).nnz == 0
    assert np.allclose(A.diag, 12.0)
    assert A.nnz == 7 * 120 - 2 * (5 * 6 + 4 * 6 + 4 * 5)


def test_graph_laplacian_properties():
    A = workloads.graph_laplacian(5000, seed=9)
    M = A.to_scipy()
    assert (abs(M - M.T) > 0).nnz == 0                      # symmetric
    off = M - sp.diags(M.diagonal())
    v = off.data
    assert np.all(v.astype(np.float32).astype(np.float64) == v)  # fp32-exact off-diagonal values
    assert np.all((-v >= 0.1) & (-v < 1.0))
    rowsum = np.asarray(M.sum(axis=1)).ravel()
    assert np.allclose(rowsum, 1e-2, rtol=0, atol=1e-9)     # L + shift*I: row sums equal the shift
    assert np.array_equal(A.diag, M.diagonal())
    for i in range(A.n):                                    # ascending columns, diagonal stored
        c = A.col[A.rowptr[i]:A.rowptr[i + 1]]
        assert np.all(np.diff(c) > 0) and i in c
    deg = np.diff(A.rowptr) - 1
    assert 10 < deg.mean() < 40


def test_graph_deterministic():
    a = workloads.graph_laplacian(3000, seed=4)
    b = workloads.graph_laplacian(3000, seed=4)
    assert np.array_equal(a.col, b.col) and np.array_equal(a.val, b.val)
