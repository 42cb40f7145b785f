"""Summation order of the library's row-binned SpMV kernels, restated in numpy for bit-exact GPU checks.

The oracle adds every row's products left to right, as the reference's CPU code does.  The GPU kernels do the
same for short rows and use fixed-order wavefront/workgroup reductions for long ones (a row-binned SpMV), so a
long row's sum is rounded differently from the oracle's while staying deterministic.  These functions give the
exact double the kernels produce for every row, so the tests can hold the GPU to its own order bit for bit and
to the oracle within the rounding bound of recursive summation (`sum_bound`).

Tile kernel (row tiles, kernels.hpp): a row of at most WAVE_MIN entries is summed by one lane left to right
from s0 (dsplit[i]*x[i], or 0); a row of WAVE_MIN+1 .. TILE_PLAN_NNZ entries by a wavefront (lane l adds
products l, l+64, ... in order, then the xor butterfly over the 64 lanes) and the total added to s0; a longer
row by the whole 256-lane workgroup (lane t adds products t, t+256, ...; xor butterfly per wavefront; the four
wavefront totals added in order), then s0 + total when dsplit is given.

Panel kernel (panel-swept layout, jpcg.hip): a row's runs inside each staged chunk are added in layout order;
a run of fewer than RUN_SCAN entries product by product, a longer one as one wavefront total.
"""
import numpy as np

WAVE_MIN = 64          # kWaveRowMinHost
TILE_PLAN_NNZ = 2044   # kTilePlanNnz
BLOCK = 256            # threads of a tile workgroup
RUN_SCAN = 15          # kPanelRunScan
EPS = np.finfo(np.float64).eps / 2


def butterfly(v):
    """xor butterfly over 64 lane values: v[l] += v[l ^ o] for o = 32 .. 1; every lane ends with v[0]."""
    v = np.array(v, dtype=np.float64)
    lanes = np.arange(64)
    for o in (32, 16, 8, 4, 2, 1):
        v = v + v[lanes ^ o]
    return v[0]


def strided_partials(q, lanes):
    """Lane l's left-to-right sum of q[l], q[l+lanes], ... (0.0 for a lane with none)."""
    p = np.zeros(lanes)
    for s in range(0, q.size, lanes):
        c = q[s:s + lanes]
        p[:c.size] = p[:c.size] + c
    return p


def wave_total(q):
    return butterfly(strided_partials(np.asarray(q, dtype=np.float64), 64))


def block_total(q):
    p = strided_partials(np.asarray(q, dtype=np.float64), BLOCK)
    w = [butterfly(p[64 * i:64 * (i + 1)]) for i in range(BLOCK // 64)]
    s = w[0]
    for t in w[1:]:
        s = s + t
    return s


def products(rp, col, val, x):
    return np.asarray(val, dtype=np.float64) * np.asarray(x, dtype=np.float64)[col]


def tile_spmv(rp, col, val, x, dsplit=None):
    """y = A x in the tile kernel's order (rows of any length)."""
    n = rp.size - 1
    q = products(rp, col, val, x)
    lens = np.diff(rp)
    rows = np.repeat(np.arange(n), lens)
    short = lens[rows] <= WAVE_MIN
    y = np.zeros(n) if dsplit is None else np.asarray(dsplit, dtype=np.float64) * np.asarray(x, dtype=np.float64)
    np.add.at(y, rows[short], q[short])          # unbuffered, in index order: left to right per row
    for i in np.nonzero(lens > WAVE_MIN)[0]:
        seg = q[rp[i]:rp[i + 1]]
        if lens[i] <= TILE_PLAN_NNZ:
            y[i] = y[i] + wave_total(seg)
        else:
            t = block_total(seg)
            y[i] = t if dsplit is None else y[i] + t
    return y


def panel_spmv(plan, n, val, x, dsplit=None):
    """y = A x in the panel kernel's order, from the library's panel plan (device.panel_plan) of the CSR whose
    values are val (CSR order)."""
    seg, perm, P, br = plan["seg"], plan["perm"], plan["P"], plan["block_rows"]
    word = plan["lrow"].astype(np.int64)
    nnz = perm.size
    q_seg = np.repeat(np.arange(seg.size - 1), np.diff(seg))
    rows = (q_seg // P) * br + (word & 0xFFF)
    cols = (q_seg % P) * 65536 + plan["off"].astype(np.int64)
    q = np.asarray(val, dtype=np.float64)[perm] * np.asarray(x, dtype=np.float64)[cols]
    code = word >> 12
    start = np.nonzero(code > 0)[0]
    ends = np.append(start[1:], nnz)
    keep = np.ones(nnz, bool)
    for s, e in zip(start[code[start] >= RUN_SCAN], ends[code[start] >= RUN_SCAN]):
        q[s] = wave_total(q[s:e])                # the run's total stands in for its first entry
        keep[s + 1:e] = False
    y = np.zeros(n) if dsplit is None else np.asarray(dsplit, dtype=np.float64) * np.asarray(x, dtype=np.float64)
    np.add.at(y, rows[keep], q[keep])
    return y


def sum_bound(rp, col, val, x, dsplit=None):
    """Per-row bound on |computed - exact| for any summation order of the row's products (and s0):
    (L+1) * u * sum |terms|, the recursive-summation bound with L+1 terms; two orders differ by at most twice it."""
    n = rp.size - 1
    lens = np.diff(rp)
    rows = np.repeat(np.arange(n), lens)
    a = np.zeros(n)
    np.add.at(a, rows, np.abs(products(rp, col, val, x)))
    if dsplit is not None:
        a += np.abs(np.asarray(dsplit) * np.asarray(x))
    return (lens + 2) * EPS * a * 1.01
