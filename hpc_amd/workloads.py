"""Seeded synthetic workloads (host numpy arrays), built by the C++ generators in libxjpcg.so.

`draw`/`uniform` mirror the counter-based SplitMix64 stream of
hpc_amd/csrc/xjpcg_rng.h in numpy, so tests can check the C generators
value for value.  Matrices come back as CSR numpy arrays (int32 rowptr/col,
fp64 values, ascending columns, diagonal stored).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import lib, ptr

GOLDEN = np.uint64(0x9E3779B97F4A7C15)

STENCIL_2D5, STENCIL_3D7, STENCIL_3D27 = 0, 1, 2


def _mix64(z: np.ndarray) -> np.ndarray:
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def draw(seed: int, stream: int, i) -> np.ndarray:
    """numpy mirror of xjpcg_draw(seed, stream, i)."""
    with np.errstate(over="ignore"):
        key = _mix64(np.uint64(seed) + GOLDEN * np.uint64(stream + 1))
        return _mix64(key + GOLDEN * (np.asarray(i, dtype=np.uint64) + np.uint64(1)))


def u01(r: np.ndarray) -> np.ndarray:
    return (r >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def uniform_np(seed: int, stream: int, n: int, lo: float, hi: float) -> np.ndarray:
    return lo + (hi - lo) * u01(draw(seed, stream, np.arange(n, dtype=np.uint64)))


def uniform(seed: int, stream: int, n: int, lo: float, hi: float) -> np.ndarray:
    """C generator: out[i] = lo + (hi-lo)*u01(draw(seed, stream, i))."""
    out = np.empty(n, dtype=np.float64)
    lib().xjpcg_fill_uniform(seed, stream, n, lo, hi, out.ctypes.data)
    return out


@dataclass
class CsrMatrix:
    rowptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    diag: np.ndarray
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.rowptr.size - 1

    @property
    def nnz(self) -> int:
        return self.col.size

    def matvec(self, x: np.ndarray) -> np.ndarray:
        """Host y = A x, rows summed left to right in fp64 (used to build right-hand sides)."""
        y = np.empty(self.n, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        lib().xjpcg_host_csr_matvec(self.n, ptr(self.rowptr), ptr(self.col), ptr(self.val), ptr(x), ptr(y))
        return y

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.n, self.n))


def stencil(kind: int, nx: int, ny: int, nz: int, coef) -> CsrMatrix:
    L = lib()
    nnz = L.xjpcg_stencil_nnz(kind, nx, ny, nz)
    if nnz < 0:
        raise ValueError("bad stencil arguments")
    n = nx * ny * nz
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    rp = np.empty(n + 1, dtype=np.int32)
    col = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    diag = np.empty(n, dtype=np.float64)
    if L.xjpcg_stencil_fill(kind, nx, ny, nz, ptr(coef), ptr(rp), ptr(col), ptr(val), ptr(diag)) != 0:
        raise ValueError("stencil generation failed")
    return CsrMatrix(rp, col, val, diag, {"kind": kind, "grid": (nx, ny, nz), "coef": coef.tolist()})


def laplace2d(m: int) -> CsrMatrix:
    """2D 5-point Dirichlet Laplacian on an m x m grid: diag 4, neighbours -1."""
    return stencil(STENCIL_2D5, m, m, 1, [1.0, 1.0])


def stencil27(m: int) -> CsrMatrix:
    """3D 27-point operator on an m^3 grid: diag 26, all neighbours -1."""
    return stencil(STENCIL_3D27, m, m, m, [26.0, -1.0])


def graph_laplacian(n: int, mu: float = 2.5, sigma: float = 1.0, cap: int = 2000, seed: int = 2,
                    wlo: float = 0.1, whi: float = 1.0, shift: float = 1e-2) -> CsrMatrix:
    """Symmetric random graph Laplacian + shift*I with fp32-exact off-diagonal values."""
    L = lib()
    g = L.xjpcg_graph_create(n, mu, sigma, cap, seed, wlo, whi, shift)
    if not g:
        raise ValueError("graph generation failed")
    try:
        nnz = L.xjpcg_graph_nnz(g)
        rp = np.empty(n + 1, dtype=np.int32)
        col = np.empty(nnz, dtype=np.int32)
        val = np.empty(nnz, dtype=np.float64)
        diag = np.empty(n, dtype=np.float64)
        L.xjpcg_graph_copy(g, ptr(rp), ptr(col), ptr(val), ptr(diag))
    finally:
        L.xjpcg_graph_free(g)
    return CsrMatrix(rp, col, val, diag, {"kind": "graph", "n": n, "mu": mu, "sigma": sigma, "cap": cap,
                                          "seed": seed, "shift": shift})


# ---------------------------------------------------------------------------
# named workloads


@dataclass
class System:
    A: CsrMatrix
    b: np.ndarray
    tol: float
    max_iter: int
    x_true: np.ndarray | None = None
    name: str = ""


def c1_laplace2d(m: int = 256, seed: int = 0) -> System:
    """2D 5-pt Laplacian m^2 (default 256^2): x_true ~ U[-1,1] (seed 0, stream 0), b = A x_true;
    stop at 1e-12 or 5000 iterations."""
    A = laplace2d(m)
    xt = uniform(seed, 0, A.n, -1.0, 1.0)
    return System(A, A.matvec(xt), 1e-12, 5000, xt, f"laplace2d_{m}")


def c2_stencil27(m: int = 160, seed: int = 1) -> System:
    """3D 27-pt m^3 (default 160^3): b = A*1 + U[-1e-3, 1e-3] (seed 1, stream 0); tol 1e-10."""
    A = stencil27(m)
    b = A.matvec(np.ones(A.n)) + uniform(seed, 0, A.n, -1e-3, 1e-3)
    return System(A, b, 1e-10, 20000, None, f"stencil27_{m}")


def c3_graph(n: int = 4_000_000, seed: int = 2) -> System:
    """Graph Laplacian + 1e-2 I (lognormal degrees mu 2.5 sigma 1 cap 2000, weights U[0.1,1] in fp32);
    b ~ U[-1,1] (seed, stream 4); tol 1e-10."""
    A = graph_laplacian(n, seed=seed)
    b = uniform(seed, 4, n, -1.0, 1.0)
    return System(A, b, 1e-10, 100000, None, f"graph_{n}")


def gaussian_random_field(m: int, length: float, seed: int) -> np.ndarray:
    """Stationary Gaussian random field on an m x m periodic grid, squared-exponential covariance
    exp(-d^2 / (2 length^2)), variance 1: white noise (numpy PCG64 seeded with `seed`) filtered in
    Fourier space by the square root of the covariance spectrum, then scaled to unit sample variance."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((m, m))
    k = np.fft.fftfreq(m) * 2 * np.pi
    kk = k[:, None] ** 2 + k[None, :] ** 2
    spec = np.exp(-0.5 * kk * length ** 2)  # spectrum of the squared-exponential kernel (up to a constant)
    g = np.real(np.fft.ifft2(np.fft.fft2(w) * np.sqrt(spec)))
    g -= g.mean()
    return g / g.std()


def diffusion2d(k_face_x: np.ndarray, k_face_y: np.ndarray) -> CsrMatrix:
    """Cell-centred 2D diffusion on an m x m grid with Dirichlet walls.

    k_face_x[i, j] is the conductivity of the face on the -x side of cell (row i, column j) for
    j = 0..m (shape m x (m+1)); k_face_y likewise on the -y side (shape (m+1) x m).  Interior face
    between cells a, b with conductivity k adds k to A_aa and A_bb and -k to A_ab, A_ba; a wall
    face adds its k to the adjacent cell's diagonal.  Rows are i*m + j, columns ascending."""
    m = k_face_x.shape[0]
    n = m * m
    diag = (k_face_x[:, :-1] + k_face_x[:, 1:] + k_face_y[:-1, :] + k_face_y[1:, :]).ravel()
    rows, cols, vals = [np.arange(n)], [np.arange(n)], [diag]
    idx = np.arange(n).reshape(m, m)
    # x faces between (i, j-1) and (i, j), j = 1..m-1
    a, b = idx[:, :-1].ravel(), idx[:, 1:].ravel()
    kx = k_face_x[:, 1:-1].ravel()
    ay, by = idx[:-1, :].ravel(), idx[1:, :].ravel()
    ky = k_face_y[1:-1, :].ravel()
    rows += [a, b, ay, by]
    cols += [b, a, by, ay]
    vals += [-kx, -kx, -ky, -ky]
    import scipy.sparse as sp
    M = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    M.sort_indices()
    return CsrMatrix(M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.float64), diag.copy(),
                     {"kind": "diffusion2d", "m": m})


def c5_diffusion(m: int = 512, length: float = 16.0, seed: int = 4):
    """2D variable-coefficient diffusion m^2 (default 512^2): face conductivities exp(GRF) with
    GRF length scale 16 cells, variance 1 (seed 4); b ~ U[-1,1] (seed 4, stream 1); x_target ~ U[-1,1]
    (seed 4, stream 2).  Returns (CsrMatrix, b, x_target)."""
    g = gaussian_random_field(m, length, seed)
    gx = np.empty((m, m + 1))
    gx[:, 1:-1] = 0.5 * (g[:, :-1] + g[:, 1:])
    gx[:, 0], gx[:, -1] = g[:, 0], g[:, -1]
    gy = np.empty((m + 1, m))
    gy[1:-1, :] = 0.5 * (g[:-1, :] + g[1:, :])
    gy[0, :], gy[-1, :] = g[0, :], g[-1, :]
    A = diffusion2d(np.exp(gx), np.exp(gy))
    b = uniform(seed, 1, A.n, -1.0, 1.0)
    xt = uniform(seed, 2, A.n, -1.0, 1.0)
    return A, b, xt


def c4_systems(count: int = 512, nmin: int = 5000, nmax: int = 400_000, nrhs: int = 16, seed: int = 3):
    """Batched workload: `count` anisotropic diffusion systems (2D 5-pt or 3D 7-pt, coefficients
    exp(0.5 N(0,1)) rounded to fp32), n log-uniform in [nmin, nmax]; nrhs right-hand sides
    U[-1,1] each.  Returns (list of CsrMatrix, list of (n, nrhs) arrays)."""
    mats, rhs = [], []
    for s in range(count):
        u = u01(draw(seed, 10, np.arange(8 * s, 8 * s + 8, dtype=np.uint64)))
        ntarget = math.exp(math.log(nmin) + (math.log(nmax) - math.log(nmin)) * u[0])
        is3d = u[1] < 0.5
        z = [math.sqrt(-2.0 * math.log(1.0 - u[2 + 2 * k])) * math.cos(2 * math.pi * u[3 + 2 * k]) for k in range(3)]
        k = [float(np.float32(math.exp(0.5 * zz))) for zz in z]
        if is3d:
            m = max(2, round(ntarget ** (1.0 / 3.0)))
            A = stencil(STENCIL_3D7, m, m, max(2, round(ntarget / (m * m))), k)
        else:
            m = max(2, round(math.sqrt(ntarget)))
            A = stencil(STENCIL_2D5, m, max(2, round(ntarget / m)), 1, k[:2])
        mats.append(A)
        B = uniform(seed, 1000 + s, A.n * nrhs, -1.0, 1.0).reshape(A.n, nrhs)
        rhs.append(B)
    return mats, rhs
