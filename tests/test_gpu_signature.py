"""Signature files to a device CSR (xjpcg_sig_to_csr) and the GPU kernels on it.

The CSR built from the signature equals the one the GPU COO ingest builds from
the original matrix (same rows, columns, duplicates in input order), so the
SpMV is bit-exact against the oracle and the solve matches it."""
import os
import sys

import numpy as np
import pytest

from hpc_amd import workloads
from oracle import oracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "signature_cases.npz")


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def coo_of(A):
    return np.repeat(np.arange(A.n), np.diff(A.rowptr)).astype(np.uint32), A.col.astype(np.uint32), A.val


def test_sig_csr_equals_coo_ingest(cuda, tmp_path):
    from hpc_amd import device
    from hpc_amd.signature import Signature
    n, k = 20000, 150000
    rng = np.random.default_rng(11)
    r = rng.integers(0, n, k).astype(np.uint32)
    c = rng.integers(0, n, k).astype(np.uint32)
    r[:3000], c[:3000] = r[-3000:], c[-3000:]
    v = rng.standard_normal(k)
    Signature.from_coo(n, n, r, c, v).store(str(tmp_path))
    sig = Signature.load(str(tmp_path))
    _say(f"signature -> CSR n={n} entries={sig.info()['entries']}")
    rp, col, val = sig.to_csr(cuda)
    rp2, c2, pm = device.csr_from_coo(n, r, c, 0, cuda)
    assert np.array_equal(rp.cpu().numpy(), rp2.cpu().numpy())
    assert np.array_equal(col.cpu().numpy(), c2.cpu().numpy())
    assert np.array_equal(val.cpu().numpy(), v[pm.cpu().numpy()])


def test_sig_golden_spmv_bitexact(cuda, tmp_path):
    """The reference partitioner's files for the 20 x 20 5-point case, decoded on the GPU path."""
    import torch
    from hpc_amd import device
    from hpc_amd.signature import Signature
    with np.load(GOLDEN) as z:
        for fn in oracle.SIG_FILES:
            (tmp_path / fn).write_bytes(z[f"stencil/{fn}"].tobytes())
        n = int(z["stencil/n"])
        r, c, v = z["stencil/row"], z["stencil/col"], z["stencil/val"]
    _say(f"golden signature spmv n={n}")
    sig = Signature.load(str(tmp_path))
    rp, col, val = sig.to_csr(cuda)
    op = device.CsrOperator(rp, col, val, rowptr_host=rp.cpu().numpy())
    x = np.random.default_rng(2).uniform(-1, 1, n)
    y = op.spmv(torch.from_numpy(x).to(cuda)).cpu().numpy()
    o = np.lexsort((c, r))  # rows, then columns ascending, ties in input order: the CSR order
    orp, ocol, oval, _ = oracle.coo_to_csr(n, r[o], c[o], v[o])
    assert np.array_equal(col.cpu().numpy(), ocol)
    assert np.array_equal(y, oracle.spmv(orp, ocol, oval, x))


def test_sig_solve_c1(cuda, tmp_path):
    """C1 (2D 5-point 256^2) through signature files: same iterations and solution as the oracle."""
    import torch
    from hpc_amd import device
    from hpc_amd.signature import Signature
    s = workloads.c1_laplace2d(256)
    A = s.A
    r, c, v = coo_of(A)
    Signature.from_coo(A.n, A.n, r, c, v).store(str(tmp_path))
    sig = Signature.load(str(tmp_path))
    _say(f"signature C1 n={A.n} {sig.info()}")
    rp, col, val = sig.to_csr(cuda)
    assert np.array_equal(rp.cpu().numpy(), A.rowptr) and np.array_equal(col.cpu().numpy(), A.col)
    op = device.CsrOperator(rp, col, val, rowptr_host=A.rowptr)
    x, info = op.solve(torch.from_numpy(A.diag).to(cuda), torch.from_numpy(s.b).to(cuda), s.max_iter, s.tol)
    ref = oracle.jpcg(A.rowptr, A.col, A.val, A.diag, s.b, s.max_iter, s.tol)
    x = x.cpu().numpy()
    _say(f"  gpu executed={info.executed} oracle={ref['executed']}")
    assert abs(info.executed - ref["executed"]) <= 1
    assert np.linalg.norm(x - ref["x"]) <= 1e-10 * np.linalg.norm(ref["x"])
