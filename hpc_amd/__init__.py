"""MI355X-native xJPCG: Jacobi-preconditioned CG and fp64 CSR SpMV in HIP for gfx950.

Layers:
  pcg      -- Python mirror of the pcg.h handle API (host arrays, status codes)
  device   -- torch-tensor operators over the device-level C ABI (SpMV, solve, mixed, SDDMM)
  batched  -- many uneven systems x several right-hand sides in one persistent launch
  autograd -- torch.autograd.Function through the solve, adjoint JPCG + SDDMM backward
  workloads-- seeded synthetic matrices and vectors
"""
from ._lib import LIB_PATH, lib  # noqa: F401

__all__ = ["LIB_PATH", "lib"]
