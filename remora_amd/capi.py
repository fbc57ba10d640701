"""ctypes binding of ``include/remora_hip/capi.h`` (libremora_hip.so).

This is the Python-side mirror of the C-ABI that Remora's hip_tag kernels
bind to.  It loads only the in-tree HIP library and raises if it is missing:
there is no CPU fallback.  torch is imported first when available so that the
library and torch share one HIP runtime in the process.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libremora_hip.so")

SUCCESS = 0
INVALID_VALUE = 1
INVALID_HANDLE = 2
HIP_ERROR = 3
ALLOC_FAILED = 4
NOT_POSITIVE_DEFINITE = 5
NOT_CONVERGED = 6
NOT_SUPPORTED = 7

FACTOR_FP32 = 0
FACTOR_BF16X3 = 1


class BatchConfig(C.Structure):
    """rmh_batch_config."""
    _fields_ = [("nb", C.c_int), ("wg_per_cu", C.c_int), ("lpt_order", C.c_int)]


class RemoraError(RuntimeError):
    """A non-success rmh_status, with rmh_last_error()'s detail."""

    def __init__(self, status: int, detail: str):
        super().__init__(f"rmh status {status}: {detail}")
        self.status = status
        self.detail = detail


_lib = None

# name -> (restype, argtypes) for every function the header declares.
P, S, I, D, F, I64 = C.c_void_p, C.c_size_t, C.c_int, C.c_double, C.c_float, C.c_int64
ST = C.c_int  # rmh_status
SIGNATURES = {
    "rmh_create": (ST, [C.POINTER(C.c_void_p), I]),
    "rmh_destroy": (ST, [P]),
    "rmh_status_string": (C.c_char_p, [ST]),
    "rmh_last_error": (C.c_char_p, []),
    "rmh_dgemm": (ST, [P, I, I, S, S, S, D, P, S, P, S, D, P, S, P]),
    "rmh_sgemm": (ST, [P, I, I, S, S, S, F, P, S, P, S, F, P, S, P]),
    "rmh_dsyrk": (ST, [P, I, I, S, S, D, P, S, D, P, S, P]),
    "rmh_ssyrk": (ST, [P, I, I, S, S, F, P, S, F, P, S, P]),
    "rmh_dtrsm": (ST, [P, I, I, I, I, S, S, D, P, S, P, S, P]),
    "rmh_strsm": (ST, [P, I, I, I, I, S, S, F, P, S, P, S, P]),
    "rmh_dtrsv": (ST, [P, I, I, I, S, P, S, P, S, P]),
    "rmh_strsv": (ST, [P, I, I, I, S, P, S, P, S, P]),
    "rmh_dcopy_matrix": (ST, [P, I, S, S, P, S, P, S, P]),
    "rmh_scopy_matrix": (ST, [P, I, S, S, P, S, P, S, P]),
    "rmh_dtrmv": (ST, [P, I, I, I, S, P, S, P, S, P]),
    "rmh_strmv": (ST, [P, I, I, I, S, P, S, P, S, P]),
    "rmh_dtrmm": (ST, [P, I, I, I, I, S, S, D, P, S, P, S, P]),
    "rmh_strmm": (ST, [P, I, I, I, I, S, S, F, P, S, P, S, P]),
    "rmh_dgemv": (ST, [P, I, S, S, D, P, S, P, S, D, P, S, P]),
    "rmh_sgemv": (ST, [P, I, S, S, F, P, S, P, S, F, P, S, P]),
    "rmh_dpotrf": (ST, [P, I, S, P, S, P, P]),
    "rmh_spotrf": (ST, [P, I, S, P, S, P, P]),
    "rmh_dpotrf_noalloc": (ST, [P, I, S, P, S, P, P]),
    "rmh_dpotrs": (ST, [P, I, S, S, P, S, P, S, P]),
    "rmh_spotrs": (ST, [P, I, S, S, P, S, P, S, P]),
    "rmh_dsposv": (ST, [P, I, I, S, S, P, S, P, S, P, S, D, I, C.POINTER(I), C.POINTER(D), P]),
    "rmh_dlaswp": (ST, [P, I, S, S, P, S, I, P, P]),
    "rmh_slaswp": (ST, [P, I, S, S, P, S, I, P, P]),
    "rmh_dpstrf": (ST, [P, I, S, P, S, P, P, P]),
    "rmh_spstrf": (ST, [P, I, S, P, S, P, P, P]),
    "rmh_dchol_update": (ST, [P, I, S, D, D, P, S, P, S, P, P]),
    "rmh_schol_update": (ST, [P, I, S, F, F, P, S, P, S, P, P]),
    "rmh_dpotrf_adjoint": (ST, [P, I, S, P, S, P, S, P, S, P]),
    "rmh_spotrf_adjoint": (ST, [P, I, S, P, S, P, S, P, S, P]),
    "rmh_dposv_batched": (ST, [P, S, I, P, P, P, S, P, P, P, P, P, P]),
    "rmh_sposv_batched": (ST, [P, S, I, P, P, P, S, P, P, P, P, P, P]),
    "rmh_dpotrf_batched": (ST, [P, S, I, P, P, P, P, P, P, P]),
    "rmh_spotrf_batched": (ST, [P, S, I, P, P, P, P, P, P, P]),
    "rmh_dpotrs_batched": (ST, [P, S, I, P, P, P, S, P, P, P, P, P]),
    "rmh_spotrs_batched": (ST, [P, S, I, P, P, P, S, P, P, P, P, P]),
    "rmh_batched_lpt_order": (ST, [S, P, P]),
    "rmh_batched_autotune": (ST, [P, I, S, P, S, P, C.POINTER(D)]),
    "rmh_set_potrf_blocking": (ST, [P, S, S]),
    "rmh_set_potrf_leaf": (ST, [P, S]),
    "rmh_set_potrf_reserve": (ST, [P, I, S, S]),
    "rmh_set_potrf_lookahead": (ST, [P, I]),
    "rmh_set_spotrf_precision": (ST, [P, I]),
    "rmh_profile_enable": (ST, [P, I]),
    "rmh_profile_reset": (ST, [P]),
    "rmh_profile_read": (ST, [P, C.c_char_p, C.POINTER(D), C.POINTER(D), C.POINTER(I64)]),
}


def load(path: str | None = None):
    """Load libremora_hip.so and declare every exported signature."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    try:  # share torch's HIP runtime when torch is present
        import torch  # noqa: F401
    except ImportError:
        pass
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libremora_hip.so not built ({p}); run __graft_entry__.build()")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(status: int) -> None:
    if status != SUCCESS:
        lib = load()
        raise RemoraError(status, lib.rmh_last_error().decode())


class Handle:
    """RAII owner of an rmh_handle (one per thread, like a BLAS handle)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.rmh_create(C.byref(h), device))
        self.h = h

    def close(self):
        if self.h:
            check(self.lib.rmh_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def value(self):
        return self.h
