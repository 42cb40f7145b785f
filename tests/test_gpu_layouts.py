"""SpMV layouts on the GPU.  Panel-swept copy: SpMV bit-exact against the oracle (every row, long ones
included), solve parity, value refresh, split-diagonal fp32 operators.  Column-offset pattern tiles: SpMV
bit-exact against the oracle and solves bit-identical to the plain row tiles.  Diagonals: SpMV bit-exact
against the oracle, solves deterministic and at oracle parity."""
import sys

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


def scattered_csr(n, seed, max_deg=40):
    rng = np.random.default_rng(seed)
    deg = rng.integers(0, max_deg, n)
    deg[rng.random(n) < 0.001] = 3000
    rows = np.repeat(np.arange(n), deg)
    cols = rng.integers(0, n, rows.size)
    vals = rng.standard_normal(rows.size)
    o = np.lexsort((cols, rows))
    rp, col, val, _ = oracle.coo_to_csr(n, rows[o], cols[o], vals[o])
    return rp, col, val


def test_panel_spmv_bitexact_all_rows(cuda):
    from hpc_amd import device
    n = 300_000
    rp, col, val = scattered_csr(n, 12)
    op = device.CsrOperator(_t(cuda, rp), _t(cuda, col), _t(cuda, val), rowptr_host=rp)
    _say(f"n={n} nnz={col.size} creation-time layout {op.layout}")
    assert op.set_layout(2) == 2
    xv = workloads.uniform(5, 3, n, -1, 1)
    x = _t(cuda, xv)
    ref = oracle.spmv(rp, col, val, xv, threads=8)
    bound = 2 * sum_order.sum_bound(rp, col, val, xv)
    y = op.spmv(x).cpu().numpy()
    # runs of a row inside a staged chunk: short ones left to right, long ones (the 3000-entry rows) by a
    # wavefront in its fixed order
    assert np.array_equal(y, sum_order.panel_spmv(device.panel_plan(rp, col), n, val, xv))
    assert np.array_equal(y[np.diff(rp) < sum_order.RUN_SCAN], ref[np.diff(rp) < sum_order.RUN_SCAN])
    assert np.all(np.abs(y - ref) <= bound)
    # residual on the same layout
    bv = workloads.uniform(5, 4, n, -1, 1)
    r, rr = op.residual(_t(cuda, bv), x)
    assert np.array_equal(r.cpu().numpy(), bv - y)
    # tiles: left to right on rows up to the wavefront bin, each bin in its own order
    assert op.set_layout(1) == 1
    yt = op.spmv(x).cpu().numpy()
    short = np.diff(rp) <= sum_order.WAVE_MIN
    assert np.array_equal(yt[short], ref[short])
    assert np.array_equal(yt, sum_order.tile_spmv(rp, col, val, xv))
    assert np.all(np.abs(yt - ref) <= bound)


def test_panel_spmv_every_run_length(cuda):
    """Rows with 0..99 entries in one panel: every run length the row words encode (1..14, left to right),
    the wavefront path (15 and more: ends at every lane of a 64-entry step) and runs cut by a chunk's end,
    bit-exact against the kernel's order."""
    from hpc_amd import device
    n = 140_000
    rng = np.random.default_rng(21)
    deg = np.arange(n) % 100
    rows = np.repeat(np.arange(n), deg)
    cols = rng.integers(0, 65536, rows.size)
    vals = rng.standard_normal(rows.size)
    o = np.lexsort((cols, rows))
    rp, col, val, _ = oracle.coo_to_csr(n, rows[o], cols[o], vals[o])
    op = device.CsrOperator(_t(cuda, rp), _t(cuda, col), _t(cuda, val), rowptr_host=rp)
    assert op.set_layout(2) == 2
    xv = workloads.uniform(5, 7, n, -1, 1)
    y = op.spmv(_t(cuda, xv)).cpu().numpy()
    ref = oracle.spmv(rp, col, val, xv, threads=8)
    assert np.array_equal(y, sum_order.panel_spmv(device.panel_plan(rp, col), n, val, xv))
    assert np.array_equal(y[deg < sum_order.RUN_SCAN], ref[deg < sum_order.RUN_SCAN])
    assert np.all(np.abs(y - ref) <= 2 * sum_order.sum_bound(rp, col, val, xv))


def test_panel_solve_matches_oracle(cuda):
    from hpc_amd import device
    A = workloads.graph_laplacian(300_000, seed=7)
    b = workloads.uniform(7, 1, A.n, -1, 1)
    op = device.CsrOperator(_t(cuda, A.rowptr), _t(cuda, A.col), _t(cuda, A.val), rowptr_host=A.rowptr)
    diag, bd = _t(cuda, A.diag), _t(cuda, b)
    tol, maxit = 1e-10, 20000
    ref = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, b, maxit, tol, threads=8)
    out = {}
    for layout in (1, 2):
        assert op.set_layout(layout) == layout
        x, info = op.solve(diag, bd, maxit, tol)
        xh = x.cpu().numpy()
        e = np.linalg.norm(xh - ref["x"]) / np.linalg.norm(ref["x"])
        _say(f"layout {layout}: {info.executed} iterations (oracle {ref['executed']}), rel err {e:.2e}")
        assert abs(info.executed - ref["executed"]) <= 1
        assert e <= 1e-10 and info.residual < tol
        out[layout] = xh
    x2, _ = op.solve(diag, bd, maxit, tol)
    assert np.array_equal(x2.cpu().numpy(), out[2])   # deterministic on the panel layout


def test_panel_value_refresh(cuda):
    """set_values with a new array and after an in-place change reaches the panel copy."""
    from hpc_amd import device
    n = 280_000
    rp, col, val = scattered_csr(n, 13, max_deg=12)
    v = _t(cuda, val)
    op = device.CsrOperator(_t(cuda, rp), _t(cuda, col), v, rowptr_host=rp)
    assert op.set_layout(2) == 2
    xv = workloads.uniform(6, 1, n, -1, 1)
    x = _t(cuda, xv)
    y1 = op.spmv(x).cpu().numpy()
    v2 = v * 2.0
    op.set_values(v2)
    assert np.array_equal(op.spmv(x).cpu().numpy(), 2.0 * y1)
    v2.mul_(0.25)                               # in place, then announce it
    op.set_values(v2)
    assert np.array_equal(op.spmv(x).cpu().numpy(), 0.5 * y1)


def test_panel_split_fp32_and_mixed(cuda):
    """fp32 off-diagonal values + fp64 split diagonal on the panel layout: same SpMV as the tiles (rows are
    short), and the mixed refined solve reaches the tolerance."""
    import torch
    from hpc_amd import device
    A = workloads.graph_laplacian(300_000, seed=8)
    b = workloads.uniform(8, 1, A.n, -1, 1)
    srp, scol, sv32, sdiag, exact = device.split_fp32(A.rowptr, A.col, A.val)
    assert exact
    inner = device.CsrOperator(_t(cuda, srp), _t(cuda, scol), _t(cuda, sv32), dsplit=_t(cuda, sdiag), rowptr_host=srp)
    outer = device.CsrOperator(_t(cuda, srp), _t(cuda, scol), _t(cuda, sv32), dsplit=_t(cuda, sdiag), rowptr_host=srp)
    x = _t(cuda, workloads.uniform(8, 2, A.n, -1, 1))
    assert inner.set_layout(1) == 1
    yt = inner.spmv(x).clone()
    assert inner.set_layout(2) == 2
    yp = inner.spmv(x)
    # each against its kernel's order (the graph's rows reach the wavefront bins of both layouts)
    xh = x.cpu().numpy()
    assert np.array_equal(yt.cpu().numpy(), sum_order.tile_spmv(srp, scol, sv32, xh, dsplit=sdiag))
    assert np.array_equal(yp.cpu().numpy(),
                          sum_order.panel_spmv(device.panel_plan(srp, scol), A.n, sv32, xh, dsplit=sdiag))
    short = torch.from_numpy(np.diff(srp) < sum_order.RUN_SCAN).to(cuda)
    assert bool((yt[short] == yp[short]).all())
    assert outer.set_layout(2) == 2
    xm, info = device.solve_mixed(outer, inner, _t(cuda, A.diag), _t(cuda, b), 1e-10)
    true = np.linalg.norm(b - A.to_scipy() @ xm.cpu().numpy()) / np.linalg.norm(b)
    _say(f"mixed on panels: outer {info.outer} inner {info.inner_total} true res {true:.2e}")
    assert true < 1e-10
    # fp32 work vectors on the panels: the refined solve still reaches 1e-10 and agrees with the fp64
    # oracle's x; a plain fp32-vector solve to 1e-6 stays within 2% of the oracle's iteration count
    inner.set_vectors(32)
    assert inner.vector_bits == 32
    xw, iw = device.solve_mixed(outer, inner, _t(cuda, A.diag), _t(cuda, b), 1e-10)
    xwh = xw.cpu().numpy()
    truew = np.linalg.norm(b - A.to_scipy() @ xwh) / np.linalg.norm(b)
    o = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, b, 20000, 1e-10, threads=8)
    errw = np.linalg.norm(xwh - o["x"]) / np.linalg.norm(o["x"])
    _say(f"mixed on panels, fp32 vectors: outer {iw.outer} inner {iw.inner_total} true res {truew:.2e} "
         f"vs oracle {errw:.2e}")
    assert truew < 1e-10 and errw < 1e-8
    x6, i6 = inner.solve(_t(cuda, A.diag), _t(cuda, b), 20000, 1e-6)
    o6 = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, b, 20000, 1e-6, threads=8)
    t6 = np.linalg.norm(b - A.to_scipy() @ x6.cpu().numpy()) / np.linalg.norm(b)
    _say(f"panels, fp32 vectors to 1e-6: {i6.executed} iterations (oracle {o6['executed']}), true res {t6:.2e}")
    assert abs(i6.executed - o6["executed"]) <= max(2, o6["executed"] // 50)
    assert i6.residual < 1e-6 and t6 < 1.5e-6


@pytest.mark.parametrize("kind", ["2d5", "3d27"])
def test_pattern_tiles_bitexact_and_identical_solves(cuda, kind):
    from hpc_amd import device
    s = workloads.c1_laplace2d(96) if kind == "2d5" else workloads.c2_stencil27(40)
    A = s.A
    op = device.CsrOperator(_t(cuda, A.rowptr), _t(cuda, A.col), _t(cuda, A.val), rowptr_host=A.rowptr)
    _say(f"{kind}: n={A.n} creation-time layout {op.layout}")
    xv = workloads.uniform(4, 1, A.n, -1, 1)
    x = _t(cuda, xv)
    ref = oracle.spmv(A.rowptr, A.col, A.val, xv)
    diag, b = _t(cuda, A.diag), _t(cuda, s.b)
    sols = {}
    for layout in (1, 3, 4):
        assert op.set_layout(layout) == layout
        assert np.array_equal(op.spmv(x).cpu().numpy(), ref)
        r, _ = op.residual(b, x)
        assert np.array_equal(r.cpu().numpy(), s.b - ref)
        xs, info = op.solve(diag, b, s.max_iter, s.tol)
        sols[layout] = (xs.cpu().numpy(), info.executed)
    # identical SpMV and identical per-tile p.Ap partials: identical solves
    assert sols[1][1] == sols[3][1] and np.array_equal(sols[1][0], sols[3][0])
    # diagonals: the same SpMV, p.Ap partials per wavefront (128 rows for fp64) instead of per tile --
    # deterministic, oracle parity
    xs2, info2 = op.solve(diag, b, s.max_iter, s.tol)
    assert info2.executed == sols[4][1] and np.array_equal(xs2.cpu().numpy(), sols[4][0])
    o = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, s.b, s.max_iter, s.tol)
    for layout in (3, 4):
        assert abs(sols[layout][1] - o["executed"]) <= 1
        assert np.linalg.norm(sols[layout][0] - o["x"]) / np.linalg.norm(o["x"]) <= 1e-10


def test_pattern_tiles_split_fp32(cuda):
    """fp32 off-diagonal values + fp64 split diagonal: pattern tiles give the tiles' SpMV bit for bit."""
    from hpc_amd import device
    A = workloads.c2_stencil27(30).A
    srp, scol, sv32, sdiag, exact = device.split_fp32(A.rowptr, A.col, A.val)
    op = device.CsrOperator(_t(cuda, srp), _t(cuda, scol), _t(cuda, sv32), dsplit=_t(cuda, sdiag), rowptr_host=srp)
    x = _t(cuda, workloads.uniform(4, 2, A.n, -1, 1))
    assert op.set_layout(1) == 1
    y1 = op.spmv(x).clone()
    assert op.set_layout(3) == 3
    assert bool((op.spmv(x) == y1).all())
    assert op.set_layout(4) == 4
    assert bool((op.spmv(x) == y1).all())


def test_pattern_layout_refused_for_graphs(cuda):
    from hpc_amd import device
    from hpc_amd._lib import XjpcgError
    A = workloads.graph_laplacian(20000, seed=4)
    op = device.CsrOperator(_t(cuda, A.rowptr), _t(cuda, A.col), _t(cuda, A.val), rowptr_host=A.rowptr)
    with pytest.raises(XjpcgError):
        op.set_layout(3)
    with pytest.raises(XjpcgError):
        op.set_layout(4)
    assert op.layout in (1, 2)


def test_stream_copy(cuda):
    import torch
    from hpc_amd import device
    for n in (1, 2, 1001, 1 << 20):
        src = torch.arange(n, dtype=torch.float64, device=cuda)
        dst = torch.full_like(src, -1.0)
        device.stream_copy(src, dst)
        assert bool((dst == src).all())


def _band_csr(n, offsets_even, offsets_odd, seed=9):
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(n):
        offs = offsets_even if i % 2 == 0 else offsets_odd
        c = i + np.asarray(offs)
        c = c[(c >= 0) & (c < n)]
        rows.append(np.full(c.size, i))
        cols.append(np.sort(c))
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    rp = np.zeros(n + 1, np.int32)
    np.cumsum(np.bincount(rows, minlength=n), out=rp[1:])
    return rp, cols.astype(np.int32), rng.standard_normal(cols.size)


@pytest.mark.parametrize("case", ["3d7-aniso", "diffusion", "band32"])
def test_pattern_tiles_more_operators(cuda, case):
    """Pattern plans built on the GPU: SpMV bit-exact against the oracle on more stencil shapes, including a
    band of exactly 32 offsets."""
    from hpc_amd import device
    if case == "3d7-aniso":
        A = workloads.stencil(workloads.STENCIL_3D7, 33, 17, 9, [0.3, 1.7, 2.9])
        rp, col, val = A.rowptr, A.col, A.val
    elif case == "diffusion":
        A, _, _ = workloads.c5_diffusion(m=64, length=4.0, seed=2)
        rp, col, val = A.rowptr, A.col, A.val
    else:
        rp, col, val = _band_csr(6000, list(range(-16, 16)), list(range(-16, 16)))
    n = rp.size - 1
    op = device.CsrOperator(_t(cuda, rp), _t(cuda, col), _t(cuda, val), rowptr_host=rp)
    xv = workloads.uniform(8, 1, n, -1, 1)
    ref = oracle.spmv(rp, col, val, xv)
    for layout in (3, 4):
        assert op.set_layout(layout) == layout
        assert np.array_equal(op.spmv(_t(cuda, xv)).cpu().numpy(), ref)


def test_pattern_layout_refused_beyond_32_offsets(cuda):
    """Rows of 32 entries each, but 48 distinct offsets per tile: the GPU plan must refuse, as the host plan."""
    from hpc_amd import device
    from hpc_amd._lib import XjpcgError
    rp, col, val = _band_csr(4000, list(range(16)) + list(range(100, 116)), list(range(16)) + list(range(200, 216)))
    assert device.pattern_plan(rp, col) is None
    op = device.CsrOperator(_t(cuda, rp), _t(cuda, col), _t(cuda, val), rowptr_host=rp)
    with pytest.raises(XjpcgError):
        op.set_layout(3)
    with pytest.raises(XjpcgError):
        op.set_layout(4)                      # 64 offsets over the whole matrix
    xv = workloads.uniform(8, 2, rp.size - 1, -1, 1)
    assert np.array_equal(op.spmv(_t(cuda, xv)).cpu().numpy(), oracle.spmv(rp, col, val, xv))


def test_diagonals_refused_for_tile_local_patterns(cuda):
    """Every tile within 32 offsets but 60 over the whole matrix (the band drifts by one column every 1000
    rows): patterns apply, diagonals are refused."""
    from hpc_amd import device
    from hpc_amd._lib import XjpcgError
    n = 40000
    rows, cols = [], []
    for i in range(n):
        a = 4 + i // 1000
        c = i + np.array([-3, -1, 0, 1, 3] + list(range(a, a + 16)))
        c = np.sort(c[(c >= 0) & (c < n)])
        rows.append(np.full(c.size, i))
        cols.append(c)
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    rp = np.zeros(n + 1, np.int32)
    np.cumsum(np.bincount(rows, minlength=n), out=rp[1:])
    col = cols.astype(np.int32)
    val = np.random.default_rng(3).standard_normal(col.size)
    op = device.CsrOperator(_t(cuda, rp), _t(cuda, col), _t(cuda, val), rowptr_host=rp)
    assert op.set_layout(3) == 3
    with pytest.raises(XjpcgError):
        op.set_layout(4)
    assert op.layout == 3
    xv = workloads.uniform(8, 4, n, -1, 1)
    assert np.array_equal(op.spmv(_t(cuda, xv)).cpu().numpy(), oracle.spmv(rp, col, val, xv))


def test_diagonals_value_refresh_and_boundary_rows(cuda):
    """Diagonals keep their own value copy: set_values (and values changed in place) refresh it before the
    next SpMV; boundary rows with missing diagonals and a row with no entries at all."""
    import torch
    from hpc_amd import device
    A = workloads.stencil(workloads.STENCIL_3D7, 21, 13, 7, [0.5, 1.5, 2.5])
    rp, col = A.rowptr.copy(), A.col.copy()
    val = A.val.copy()
    # empty the last row
    k0, k1 = rp[-2], rp[-1]
    col = np.delete(col, np.arange(k0, k1))
    val = np.delete(val, np.arange(k0, k1))
    rp[-1] = k0
    n = rp.size - 1
    v = _t(cuda, val)
    op = device.CsrOperator(_t(cuda, rp), _t(cuda, col), v, rowptr_host=rp)
    assert op.set_layout(4) == 4
    xv = workloads.uniform(8, 3, n, -1, 1)
    x = _t(cuda, xv)
    assert np.array_equal(op.spmv(x).cpu().numpy(), oracle.spmv(rp, col, val, xv))
    v.mul_(-2.5)                              # in place, then announced
    op.set_values(v)
    assert np.array_equal(op.spmv(x).cpu().numpy(), oracle.spmv(rp, col, -2.5 * val, xv))
    w = torch.from_numpy(np.ascontiguousarray(val * 3.0)).to(cuda)
    op.set_values(w)
    assert np.array_equal(op.spmv(x).cpu().numpy(), oracle.spmv(rp, col, 3.0 * val, xv))


def test_diagonal_solve_odd_rows(cuda):
    """Diagonal-layout CG on an odd number of rows below 327,680 (one row per lane; the last workgroup runs
    partly empty).  Oracle parity and repeatability.  The two-rows-per-lane padding row is covered by the
    91 x 81 x 81 operator below."""
    from hpc_amd import device
    A = workloads.stencil(workloads.STENCIL_3D7, 21, 13, 7, [0.5, 1.5, 2.5])
    assert A.n % 2 == 1
    b = workloads.uniform(5, 2, A.n, -1, 1)
    op = device.CsrOperator(_t(cuda, A.rowptr), _t(cuda, A.col), _t(cuda, A.val), rowptr_host=A.rowptr)
    assert op.set_layout(4) == 4 and op.solve_layout == 4
    diag, bt = _t(cuda, A.diag), _t(cuda, b)
    x1, i1 = op.solve(diag, bt, 2000, 1e-12)
    x2, i2 = op.solve(diag, bt, 2000, 1e-12)
    assert i1.executed == i2.executed and bool((x1 == x2).all())
    o = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, b, 2000, 1e-12)
    _say(f"n={A.n} diagonals: {i1.executed} iterations, oracle {o['executed']}")
    assert abs(i1.executed - o["executed"]) <= 1
    assert np.linalg.norm(x1.cpu().numpy() - o["x"]) / np.linalg.norm(o["x"]) <= 1e-10


def test_equal_matrices_give_identical_solves(cuda):
    """The solve layout follows from the matrix, not from creation-time timings: two operators built from
    the same 81^3 27-point matrix (two rows per lane from 327,680 rows) both solve on the diagonals and give
    bit-identical iterates, at oracle parity."""
    from hpc_amd import device
    s = workloads.c2_stencil27(81)
    A = s.A
    assert A.n >= 5 << 16
    out = []
    for _ in range(2):
        op = device.CsrOperator(_t(cuda, A.rowptr), _t(cuda, A.col), _t(cuda, A.val), rowptr_host=A.rowptr)
        assert op.solve_layout == 4
        x, info = op.solve(_t(cuda, A.diag), _t(cuda, s.b), s.max_iter, 1e-8)
        out.append((x.cpu().numpy(), info.executed, info.residual))
        op.close()
    assert out[0][1] == out[1][1] and out[0][2] == out[1][2] and np.array_equal(out[0][0], out[1][0])
    o = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, s.b, s.max_iter, 1e-8)
    _say(f"n={A.n}: {out[0][1]} iterations, oracle {o['executed']}")
    assert abs(out[0][1] - o["executed"]) <= 1
    assert np.linalg.norm(out[0][0] - o["x"]) / np.linalg.norm(o["x"]) <= 1e-10


def test_equal_graph_matrices_give_identical_panel_solves(cuda):
    """The random-graph workload at 300,000 rows (columns scattered, no int16 stream): every operator built on
    it solves on the panel layout by rule, whatever its creation-time timings, so two operators give
    bit-identical iterates and both match the oracle."""
    from hpc_amd import device
    A = workloads.graph_laplacian(300_000, seed=9)
    b = workloads.uniform(9, 1, A.n, -1, 1)
    out = []
    for _ in range(2):
        op = device.CsrOperator(_t(cuda, A.rowptr), _t(cuda, A.col), _t(cuda, A.val), rowptr_host=A.rowptr)
        assert op.solve_layout == 2 and not op.compact
        x, info = op.solve(_t(cuda, A.diag), _t(cuda, b), 20000, 1e-10)
        out.append((x.cpu().numpy(), info.executed, info.residual))
        op.close()
    assert out[0][1] == out[1][1] and out[0][2] == out[1][2] and np.array_equal(out[0][0], out[1][0])
    o = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, b, 20000, 1e-10, threads=8)
    _say(f"graph n={A.n} on panels: {out[0][1]} iterations, oracle {o['executed']}")
    assert abs(out[0][1] - o["executed"]) <= 1
    assert np.linalg.norm(out[0][0] - o["x"]) / np.linalg.norm(o["x"]) <= 1e-10


def test_diagonal_solve_7point_above_floor(cuda):
    """An anisotropic 3D 7-point operator with an odd number of rows above 327,680: the
    operator solves on the diagonals by itself (two rows per lane, the last lane's second row padding) and
    matches the oracle."""
    from hpc_amd import device
    A = workloads.stencil(workloads.STENCIL_3D7, 91, 81, 81, [0.5, 1.5, 2.5])
    assert A.n % 2 == 1 and A.n >= 5 << 16
    b = workloads.uniform(6, 1, A.n, -1, 1)
    op = device.CsrOperator(_t(cuda, A.rowptr), _t(cuda, A.col), _t(cuda, A.val), rowptr_host=A.rowptr)
    assert op.solve_layout == 4
    x, info = op.solve(_t(cuda, A.diag), _t(cuda, b), 20000, 1e-10)
    o = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, b, 20000, 1e-10)
    _say(f"n={A.n} 7-point on diagonals: {info.executed} iterations, oracle {o['executed']}")
    assert abs(info.executed - o["executed"]) <= 1
    assert np.linalg.norm(x.cpu().numpy() - o["x"]) / np.linalg.norm(o["x"]) <= 1e-10


def test_mixed_inner_on_diagonals_above_floor(cuda):
    """Mixed precision on the same odd 7-point operator: the fp32-valued inner operator (split fp64
    diagonal) solves on the diagonals, one row per lane; the refined solve reaches the tolerance in true
    residual and agrees with the fp64 oracle's solution."""
    from hpc_amd import device
    A = workloads.stencil(workloads.STENCIL_3D7, 91, 81, 81, [0.5, 1.5, 2.5])
    b = workloads.uniform(6, 1, A.n, -1, 1)
    srp, scol, sv32, sdiag, exact = device.split_fp32(A.rowptr, A.col, A.val)
    assert exact
    inner = device.CsrOperator(_t(cuda, srp), _t(cuda, scol), _t(cuda, sv32), dsplit=_t(cuda, sdiag), rowptr_host=srp)
    outer = device.CsrOperator(_t(cuda, A.rowptr), _t(cuda, A.col), _t(cuda, A.val), rowptr_host=A.rowptr)
    assert inner.solve_layout == 4
    xm, info = device.solve_mixed(outer, inner, _t(cuda, A.diag), _t(cuda, b), 1e-10)
    xh = xm.cpu().numpy()
    true = np.linalg.norm(b - A.to_scipy() @ xh) / np.linalg.norm(b)
    o = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, b, 20000, 1e-10)
    err = np.linalg.norm(xh - o["x"]) / np.linalg.norm(o["x"])
    _say(f"mixed, inner on diagonals: outer {info.outer} inner {info.inner_total} (oracle {o['executed']}) "
         f"true res {true:.2e} vs oracle {err:.2e}")
    assert true < 1e-10
    assert err < 1e-8


def test_mixed_inner_fp32_vectors_on_diagonals(cuda):
    """fp32 work vectors (r, p, Ap and J stored fp32; x and the dots fp64) in the inner operator of the refinement
    on the odd 7-point operator: the refined solve reaches 1e-10 in true residual and agrees with the fp64
    oracle's x; a plain solve with fp32 vectors to 1e-6 stays within 2% of the oracle's iterations and its
    true residual within 1.5x of the tolerance.  The flag needs the diagonal solve layout and lapses, to
    fp64 vectors, when another solve layout is set."""
    from hpc_amd import device
    from hpc_amd._lib import XjpcgError
    A = workloads.stencil(workloads.STENCIL_3D7, 91, 81, 81, [0.5, 1.5, 2.5])
    b = workloads.uniform(6, 1, A.n, -1, 1)
    srp, scol, sv32, sdiag, exact = device.split_fp32(A.rowptr, A.col, A.val)
    inner = device.CsrOperator(_t(cuda, srp), _t(cuda, scol), _t(cuda, sv32), dsplit=_t(cuda, sdiag), rowptr_host=srp)
    outer = device.CsrOperator(_t(cuda, A.rowptr), _t(cuda, A.col), _t(cuda, A.val), rowptr_host=A.rowptr)
    assert inner.solve_layout == 4 and inner.vector_bits == 64
    inner.set_vectors(32)
    assert inner.vector_bits == 32
    diag, bt = _t(cuda, A.diag), _t(cuda, b)
    xm, info = device.solve_mixed(outer, inner, diag, bt, 1e-10)
    xh = xm.cpu().numpy()
    S = A.to_scipy()
    true = np.linalg.norm(b - S @ xh) / np.linalg.norm(b)
    o = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, b, 20000, 1e-10)
    err = np.linalg.norm(xh - o["x"]) / np.linalg.norm(o["x"])
    _say(f"mixed, fp32 vectors: outer {info.outer} inner {info.inner_total} (oracle {o['executed']}) "
         f"true res {true:.2e} vs oracle {err:.2e}")
    assert info.outer >= 1 and true < 1e-10
    assert err < 1e-8
    # a plain solve with fp32 vectors on the fp64-valued operator
    outer.set_layout(4)
    outer.set_vectors(32)
    x6, i6 = outer.solve(diag, bt, 20000, 1e-6)
    o6 = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, b, 20000, 1e-6)
    t6 = np.linalg.norm(b - S @ x6.cpu().numpy()) / np.linalg.norm(b)
    _say(f"fp32 vectors to 1e-6: {i6.executed} iterations (oracle {o6['executed']}), true res {t6:.2e}")
    assert abs(i6.executed - o6["executed"]) <= max(2, o6["executed"] // 50)
    assert i6.residual < 1e-6 and t6 < 1.5e-6
    outer.set_layout(1)
    assert outer.vector_bits == 64
    with pytest.raises(XjpcgError):
        outer.set_vectors(32)
    outer.set_layout(4)
    assert outer.vector_bits == 32
    outer.set_vectors(64)
    x64, i64 = outer.solve(diag, bt, 20000, 1e-6)
    assert i64.executed == o6["executed"] or abs(i64.executed - o6["executed"]) <= 1


def test_fp32_vectors_edge_cases(cuda):
    """fp32 work vectors on the diagonal layout keep the loop's semantics: b = 0 gives x = 0 and no
    iteration; the iteration cap runs maxIter+1 rounds and reports maxIter, like the fp64 loop and the
    oracle; the records of a capped solve count 0..K-1 with the recursive residual of each round, within
    fp32 rounding of the fp64 loop's early records."""
    import torch
    from hpc_amd import device
    A = workloads.stencil(workloads.STENCIL_3D7, 91, 81, 81, [0.5, 1.5, 2.5])
    b = workloads.uniform(6, 1, A.n, -1, 1)
    op = device.CsrOperator(_t(cuda, A.rowptr), _t(cuda, A.col), _t(cuda, A.val), rowptr_host=A.rowptr)
    op.set_layout(4)
    d, bt = _t(cuda, A.diag), _t(cuda, b)
    cap = 12
    rec64 = torch.zeros((cap + 2) * 64, dtype=torch.uint8, device=cuda)
    x64, i64 = op.solve(d, bt, cap, 1e-12, records=rec64)
    op.set_vectors(32)
    xz, iz = op.solve(d, _t(cuda, np.zeros(A.n)), cap, 1e-12)
    assert iz.executed == 0 and iz.iterations == 0 and iz.residual == 0.0
    assert not xz.cpu().numpy().any()
    rec32 = torch.zeros((cap + 2) * 64, dtype=torch.uint8, device=cuda)
    x32, i32 = op.solve(d, bt, cap, 1e-12, records=rec32)
    o = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, b, cap, 1e-12)
    assert i32.executed == i64.executed == o["executed"] == cap + 1
    assert i32.iterations == i64.iterations == o["iter"] == cap
    r32 = rec32.cpu().numpy().view(oracle.RECORD_DTYPE)
    r64 = rec64.cpu().numpy().view(oracle.RECORD_DTYPE)
    assert np.array_equal(r32["iter"][1:cap + 2], np.arange(cap + 1))
    assert np.allclose(r32["res"][1:cap + 2], r64["res"][1:cap + 2], rtol=1e-4)
    assert np.allclose(r32["alpha"][1:cap + 2], r64["alpha"][1:cap + 2], rtol=1e-4)
    _say(f"fp32 vectors capped at {cap}: res {i32.residual:.6e} vs fp64 {i64.residual:.6e}; "
         f"x rel diff {np.linalg.norm(x32.cpu().numpy() - x64.cpu().numpy()) / np.linalg.norm(x64.cpu().numpy()):.2e}")
    assert abs(i32.residual - i64.residual) <= 1e-4 * i64.residual


def test_solve_layout_chosen_separately(cuda):
    """The solves keep their own creation-time layout; set_layout(L > 0) sets both, 0 restores both."""
    from hpc_amd import device
    s = workloads.c2_stencil27(40)
    A = s.A
    op = device.CsrOperator(_t(cuda, A.rowptr), _t(cuda, A.col), _t(cuda, A.val), rowptr_host=A.rowptr)
    spmv0, solve0 = op.layout, op.solve_layout
    _say(f"creation-time layouts: spmv {spmv0} solve {solve0}")
    assert spmv0 in (1, 3, 4) and solve0 in (1, 3, 4)
    diag, b = _t(cuda, A.diag), _t(cuda, s.b)
    x_auto, info_auto = op.solve(diag, b, s.max_iter, s.tol)
    assert op.set_layout(solve0) == solve0 and op.solve_layout == solve0
    x_same, info_same = op.solve(diag, b, s.max_iter, s.tol)
    assert info_same.executed == info_auto.executed and bool((x_same == x_auto).all())
    for lay in (1, 3, 4):
        assert op.set_layout(lay) == lay and op.solve_layout == lay
    assert op.set_layout(0) == spmv0 and op.solve_layout == solve0


@pytest.mark.parametrize("m", [70, 71])
def test_diagonal_spmv_two_rows_per_lane_bitexact(cuda, m):
    """27-point m^3 operators just above 327,680 rows, where the fp64 diagonal kernel takes two rows per lane
    (71^3 is odd: the last lane's second row is padding): y = A x equals the oracle's left-to-right sums."""
    from hpc_amd import device
    A = workloads.stencil27(m)
    assert A.n >= 5 << 16 and A.n % 2 == m % 2
    op = device.CsrOperator(_t(cuda, A.rowptr), _t(cuda, A.col), _t(cuda, A.val), rowptr_host=A.rowptr)
    assert op.set_layout(4) == 4
    xv = workloads.uniform(12, m, A.n, -1, 1)
    y = op.spmv(_t(cuda, xv)).cpu().numpy()
    _say(f"diagonals {m}^3 (n={A.n}): bit-exact check")
    assert np.array_equal(y, oracle.spmv(A.rowptr, A.col, A.val, xv, threads=8))
