"""ctypes binding of libxjpcg.so (the HIP/C++ library built from hpc_amd/csrc).

The library is loaded from this package directory only.  There is no CPU
fallback: when the library is missing every entry point raises, so a product
call can never silently run somewhere other than the GPU kernels.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_uint32, c_uint64, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxjpcg.so")
LOADER_PATH = os.path.join(HERE, "libxjpcg_loader.a")

_lib = None

# name -> (restype, argtypes)
_SIGNATURES = {
    # pcg.h
    "xJPCG_getErrorString": (c_char_p, [c_int]),
    "xJPCG_createHandle": (c_int, [POINTER(c_void_p), c_char_p]),
    "xJPCG_destroyHandle": (c_int, [c_void_p]),
    "xJPCG_cscSymSolver": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_int64, c_double, POINTER(c_int64), POINTER(c_double), c_int]),
    "xJPCG_cooSolver": (c_int, [c_void_p, c_uint32, c_uint32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_uint32, c_double, POINTER(c_uint32), POINTER(c_double), c_int]),
    "xJPCG_peekAtLastStatus": (c_int, [c_void_p]),
    "xJPCG_getLastMessage": (c_char_p, [c_void_p]),
    "xJPCG_getMetrics": (c_int, [c_void_p, c_void_p]),
    # xjpcg_ext.h
    "xJPCG_getHistory": (c_int, [c_void_p, c_void_p, c_uint64, POINTER(c_uint64)]),
    "xJPCG_setCooMat": (c_int, [c_void_p, c_uint32, c_uint32, c_void_p, c_void_p, c_void_p, c_int]),
    "xJPCG_setCscSymMat": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int]),
    "xJPCG_setVec": (c_int, [c_void_p, c_uint32, c_void_p, c_void_p]),
    "xJPCG_run": (c_int, [c_void_p, c_uint32, c_double, POINTER(c_uint32), POINTER(c_double)]),
    "xJPCG_getResults": (c_int, [c_void_p, POINTER(c_uint32), POINTER(c_double)]),
    "xJPCG_getX": (c_int, [c_void_p, c_void_p, c_uint32]),
    # xjpcg_dev.h
    "xjpcg_dev_last_error": (c_char_p, []),
    "xjpcg_op_create": (c_int, [POINTER(c_void_p), c_int32, c_int64, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                c_void_p, c_void_p]),
    "xjpcg_op_destroy": (c_int, [c_void_p]),
    "xjpcg_op_set_values": (c_int, [c_void_p, c_void_p]),
    "xjpcg_op_set_compact": (c_int, [c_void_p, c_int, POINTER(c_int)]),
    "xjpcg_op_info": (c_int, [c_void_p, POINTER(c_int32), POINTER(c_int64), POINTER(c_int32), POINTER(c_int32)]),
    "xjpcg_op_spmv_bytes": (c_int, [c_void_p, c_int, POINTER(c_int64)]),
    "xjpcg_op_set_path": (c_int, [c_void_p, c_int, POINTER(c_int)]),
    "xjpcg_op_set_layout": (c_int, [c_void_p, c_int, c_void_p, POINTER(c_int)]),
    "xjpcg_op_solve_layout": (c_int, [c_void_p, POINTER(c_int)]),
    "xjpcg_op_set_vectors": (c_int, [c_void_p, c_int, POINTER(c_int)]),
    "xjpcg_panel_dims": (c_int, [c_int32, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32)]),
    "xjpcg_panel_plan": (c_int, [c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 POINTER(c_int)]),
    "xjpcg_tile_plan": (c_int, [c_int32, c_void_p, c_void_p, POINTER(c_int32)]),
    "xjpcg_stream_copy": (c_int, [c_void_p, c_void_p, c_int64, c_void_p]),
    "xjpcg_csr_from_coo": (c_int, [c_uint32, c_uint32, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                   c_void_p]),
    "xjpcg_csr_from_cscsym": (c_int, [c_uint32, c_uint32, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                      POINTER(c_int64), c_void_p]),
    "xjpcg_pattern_plan": (c_int, [c_int32, c_void_p, c_void_p, c_void_p, c_void_p, POINTER(c_int)]),
    "xjpcg_spmv": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "xjpcg_residual": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, POINTER(c_double), c_void_p]),
    "xjpcg_solve": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_double, POINTER(c_uint32),
                            POINTER(c_double), POINTER(c_int64), c_void_p, c_void_p]),
    "xjpcg_sddmm": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "xjpcg_solve_mixed": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_double, c_double,
                                  c_uint32, POINTER(c_uint32), POINTER(c_int64), POINTER(c_double), c_void_p]),
    "xjpcg_csr_solve_static": (c_int, [c_int32, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_uint32, c_double, POINTER(c_uint32), POINTER(c_double), POINTER(c_int64),
                                       c_void_p]),
    "xjpcg_batch_last_error": (c_char_p, []),
    "xjpcg_batch_create": (c_int, [POINTER(c_void_p), c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_int, c_void_p, c_void_p]),
    "xjpcg_batch_destroy": (c_int, [c_void_p]),
    "xjpcg_batch_info": (c_int, [c_void_p, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32), POINTER(c_int32),
                                 POINTER(c_int32)]),
    "xjpcg_batch_set_config": (c_int, [c_void_p, c_int32, c_int32, c_int32]),
    "xjpcg_batch_set_layout": (c_int, [c_void_p, c_int32, POINTER(c_int32), c_void_p]),
    "xjpcg_batch_set_vectors": (c_int, [c_void_p, c_int32, POINTER(c_int32)]),
    "xjpcg_batch_set_path": (c_int, [c_void_p, c_int32, c_int32, POINTER(c_int32), POINTER(c_int32)]),
    "xjpcg_batch_solve": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_double, c_void_p,
                                  c_void_p, POINTER(c_int64), c_void_p]),
    "xjpcg_batch_solve_mixed": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_double,
                                        c_double, c_uint32, c_void_p, c_void_p, POINTER(c_uint32),
                                        POINTER(c_int64), c_void_p]),
    # xjpcg_sig.h
    "xjpcg_sig_last_error": (c_char_p, []),
    "xjpcg_sig_from_coo": (c_int, [POINTER(c_void_p), c_uint32, c_uint32, c_uint32, c_void_p, c_void_p, c_void_p,
                                   c_int]),
    "xjpcg_sig_from_cscsym": (c_int, [POINTER(c_void_p), c_uint32, c_uint32, c_void_p, c_void_p, c_void_p, c_int]),
    "xjpcg_sig_update": (c_int, [c_void_p, c_void_p, c_uint64]),
    "xjpcg_sig_store": (c_int, [c_void_p, c_char_p]),
    "xjpcg_sig_load": (c_int, [POINTER(c_void_p), c_char_p]),
    "xjpcg_sig_destroy": (c_int, [c_void_p]),
    "xjpcg_sig_info": (c_int, [c_void_p, c_void_p, POINTER(c_uint64), POINTER(c_uint32), POINTER(c_uint32)]),
    "xjpcg_sig_buffer": (c_int, [c_void_p, c_int, POINTER(c_void_p), POINTER(c_uint64)]),
    "xjpcg_sig_to_coo": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint64]),
    "xjpcg_sig_to_csr": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint64, POINTER(c_uint64), c_void_p]),
    # xjpcg_workloads.h
    "xjpcg_stencil_nnz": (c_int64, [c_int, c_int, c_int, c_int]),
    "xjpcg_stencil_fill": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "xjpcg_fill_uniform": (None, [c_uint64, c_uint64, c_int64, c_double, c_double, c_void_p]),
    "xjpcg_host_csr_matvec": (None, [c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "xjpcg_graph_create": (c_void_p, [c_int32, c_double, c_double, c_int32, c_uint64, c_double, c_double,
                                      c_double]),
    "xjpcg_graph_nnz": (c_int64, [c_void_p]),
    "xjpcg_graph_copy": (None, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "xjpcg_graph_free": (None, [c_void_p]),
}


class LibraryMissing(RuntimeError):
    pass


def lib():
    """Load libxjpcg.so (once) and attach argument types; raise if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (or make -C hpc_amd/csrc) first")
    # torch ships its own HIP runtime with the same soname (libamdhip64.so.7); loading torch
    # first makes libxjpcg.so bind to that one instance, so device pointers and streams from
    # torch tensors are valid in the library.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def declared_symbols():
    return list(_SIGNATURES)


def ptr(a) -> int | None:
    """Address of a numpy array or torch tensor (None for None)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


class XjpcgError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"status {status}: {message}")
        self.status = status
        self.message = message


def check_dev(status: int):
    if status != 0:
        raise XjpcgError(status, lib().xjpcg_dev_last_error().decode(errors="replace"))
