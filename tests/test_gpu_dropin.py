"""Drop-in checks on the GPU: the reference's own C example runs unmodified against this library (linked as
libXilinxPcg.so and through the reference's dlopen loader), and the step-wise handle calls give the
one-call solver's results."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from hpc_amd import workloads

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["ref_example_linked", "ref_example_loader"])
def test_reference_example_runs(cuda, name):
    exe = os.path.join(ROOT, "oracle", "_ref", name)
    if not os.path.exists(exe):
        pytest.skip("reference example not built")
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(ROOT, "hpc_amd"))
    _say(f"running {name}")
    p = subprocess.run([exe, "0"], env=env, capture_output=True, text=True, timeout=120)
    _say(p.stdout[-800:])
    assert p.returncode == 0, p.stdout + p.stderr
    blocks = re.findall(r"Num of iterations:\s+(\d+)\s+Relative residual:\s+(\S+)", p.stdout)
    assert len(blocks) == 3
    for it, res in blocks:
        assert int(it) >= 1 and float(res) <= 1e-16   # the example's tolerance; every call returned SUCCESS


def test_step_api_matches_solver(cuda):
    from hpc_amd.pcg import JPCG, Status
    s = workloads.c1_laplace2d(64)
    A = s.A
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    with JPCG(0) as h:
        one = h.coo_solve(A.n, rows, A.col, A.val, A.diag, s.b, s.max_iter, s.tol)
        assert one.status == Status.SUCCESS
    with JPCG(0) as h:
        assert h.set_coo_mat(A.n, rows, A.col, A.val) == Status.SUCCESS
        assert h.set_vec(A.n, s.b, A.diag) == Status.SUCCESS
        st, it, res = h.run(s.max_iter, s.tol)
        assert st == Status.SUCCESS and it == one.iterations and res == one.residual
        assert h.results() == (Status.SUCCESS, it, res)
        st, x = h.get_x(A.n)
        assert st == Status.SUCCESS and np.array_equal(x, one.x)
        # values-only refresh to 2A (and the diagonal to 2 diag): every quantity scales by a power of two,
        # so the iterates are exactly those of A with x halved
        assert h.set_coo_mat(A.n, None, None, 2.0 * A.val, mode=0x01) == Status.SUCCESS
        assert h.set_vec(A.n, s.b, 2.0 * A.diag) == Status.SUCCESS
        st, it2, res2 = h.run(s.max_iter, s.tol)
        st2, x2 = h.get_x(A.n)
        _say(f"steps: {it} iterations, refreshed 2A: {it2}")
        assert st == Status.SUCCESS and it2 == it and res2 == res
        assert np.array_equal(x2, 0.5 * one.x)
        # a failing tolerance keeps the results and reports EXECUTION_FAILED
        st, it3, res3 = h.run(3, 1e-14)
        assert st == Status.EXECUTION_FAILED and it3 == 3 and res3 > 1e-14
        assert h.results() == (Status.SUCCESS, it3, res3)
