import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests through the C ABI")


def _ensure_built():
    """Build what is missing (CPU-side: hipcc cross-compiles, gcc builds the oracle)."""
    lib = os.path.join(ROOT, "hpc_amd", "libxjpcg.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "hpc_amd", "csrc")], check=True)
    if not os.path.exists(os.path.join(ROOT, "tests", "c", "example_tridiag")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "c")], check=True)
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], check=True)


_ensure_built()


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    return torch.device("cuda:0")
