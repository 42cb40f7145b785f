"""bench.py command line on the CPU: the multi-GPU request is refused with a message and a non-zero exit
before any device work; --help names the modes."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_refused():
    p = _run("--gpus", "2")
    assert p.returncode != 0
    msg = json.loads(p.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_help_lists_modes():
    p = _run("--help")
    assert p.returncode == 0
    for mode in ("spmv-profile", "graph", "batched", "autograd", "signature"):
        assert mode in p.stdout
