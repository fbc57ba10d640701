"""torch-tensor front of the C-ABI (device memory and streams come from torch).

Matrices are column-major: a 2-D CUDA tensor ``X`` of shape (rows, cols)
with ``X.stride(0) == 1`` and leading dimension ``X.stride(1)``; ``colmajor``
makes one.  Every call is asynchronous on torch's current stream unless a
stream is given.  Functions mirror the BLAS-style entry points of
``include/remora_hip/capi.h`` one to one.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import capi

_DT = {torch.float64: "d", torch.float32: "s"}


def colmajor(x: torch.Tensor) -> torch.Tensor:
    """Column-major copy (or view, if already column-major) of a 2-D tensor."""
    if x.dim() != 2:
        raise ValueError("expected a matrix")
    if x.stride(0) == 1 and x.stride(1) >= max(1, x.size(0)):
        return x
    return x.t().contiguous().t()


def empty_colmajor(rows: int, cols: int, dtype=torch.float64, device="cuda") -> torch.Tensor:
    return torch.empty((cols, rows), dtype=dtype, device=device).t()


def _ld(x: torch.Tensor) -> int:
    if x.dim() != 2:
        raise ValueError("expected a matrix")
    if x.numel() == 0:
        return max(1, x.size(0))
    if x.stride(0) != 1 and x.size(0) > 1:
        raise ValueError("matrix is not column-major (stride(0) must be 1)")
    return max(1, x.stride(1) if x.size(1) > 1 else x.size(0))


def _p(x: torch.Tensor):
    return C.c_void_p(x.data_ptr())


def _stream(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _fn(prefix: str, dtype, name: str):
    if dtype not in _DT:
        raise TypeError(f"unsupported dtype {dtype}")
    return getattr(capi.load(), f"{prefix}{_DT[dtype]}{name}")


def gemm(h, A, B, C_, alpha=1.0, beta=0.0, transA=False, transB=False, stream=None):
    """C = alpha op(A) op(B) + beta C."""
    m, n = C_.shape
    k = A.shape[0] if transA else A.shape[1]
    f = _fn("rmh_", C_.dtype, "gemm")
    capi.check(f(h.value, int(transA), int(transB), m, n, k, alpha, _p(A), _ld(A), _p(B), _ld(B), beta, _p(C_),
                 _ld(C_), _stream(stream)))
    return C_


def syrk(h, A, C_, alpha=1.0, beta=0.0, upper=False, trans=False, stream=None):
    """C = alpha op(A) op(A)^T + beta C on one triangle."""
    n = C_.shape[0]
    k = A.shape[0] if trans else A.shape[1]
    f = _fn("rmh_", C_.dtype, "syrk")
    capi.check(f(h.value, int(upper), int(trans), n, k, alpha, _p(A), _ld(A), beta, _p(C_), _ld(C_),
                 _stream(stream)))
    return C_


def trsm(h, A, B, left=True, upper=False, trans=False, unit=False, alpha=1.0, stream=None):
    """Solve op(A) X = alpha B (left) or X op(A) = alpha B (right) in place."""
    m, n = B.shape
    f = _fn("rmh_", B.dtype, "trsm")
    capi.check(f(h.value, int(left), int(upper), int(trans), int(unit), m, n, alpha, _p(A), _ld(A), _p(B), _ld(B),
                 _stream(stream)))
    return B


def trsv(h, A, x, upper=False, trans=False, unit=False, incx=1, stream=None):
    """Solve op(A) x = b in place; x is a 1-D tensor read with stride incx."""
    n = A.shape[0]
    f = _fn("rmh_", A.dtype, "trsv")
    capi.check(f(h.value, int(upper), int(trans), int(unit), n, _p(A), _ld(A), _p(x), incx, _stream(stream)))
    return x


def copy_matrix(h, A, B, trans=False, stream=None):
    """B = op(A) on the device (B column-major, m x n)."""
    m, n = B.shape
    f = _fn("rmh_", B.dtype, "copy_matrix")
    capi.check(f(h.value, int(trans), m, n, _p(A), _ld(A), _p(B), _ld(B), _stream(stream)))
    return B


def trmv(h, A, x, upper=False, trans=False, unit=False, incx=1, stream=None):
    """x = op(A) x in place, A triangular; x is a 1-D tensor read with stride incx."""
    n = A.shape[0]
    f = _fn("rmh_", A.dtype, "trmv")
    capi.check(f(h.value, int(upper), int(trans), int(unit), n, _p(A), _ld(A), _p(x), incx, _stream(stream)))
    return x


def trmm(h, A, B, left=True, upper=False, trans=False, unit=False, alpha=1.0, stream=None):
    """B = alpha op(A) B (left) or alpha B op(A) (right) in place, A triangular."""
    m, n = B.shape
    f = _fn("rmh_", B.dtype, "trmm")
    capi.check(f(h.value, int(left), int(upper), int(trans), int(unit), m, n, alpha, _p(A), _ld(A), _p(B), _ld(B),
                 _stream(stream)))
    return B


def gemv(h, A, x, y, alpha=1.0, beta=0.0, trans=False, incx=1, incy=1, stream=None):
    m, n = A.shape
    f = _fn("rmh_", A.dtype, "gemv")
    capi.check(f(h.value, int(trans), m, n, alpha, _p(A), _ld(A), _p(x), incx, beta, _p(y), incy, _stream(stream)))
    return y


def potrf(h, A, upper=False, info=None, stream=None):
    """In-place Cholesky; returns the device int64 info tensor."""
    n = A.shape[0]
    if info is None:
        info = torch.zeros(1, dtype=torch.int64, device=A.device)
    f = _fn("rmh_", A.dtype, "potrf")
    capi.check(f(h.value, int(upper), n, _p(A), _ld(A), _p(info), _stream(stream)))
    return info


def potrs(h, L, B, upper=False, stream=None):
    n, nrhs = B.shape
    f = _fn("rmh_", B.dtype, "potrs")
    capi.check(f(h.value, int(upper), n, nrhs, _p(L), _ld(L), _p(B), _ld(B), _stream(stream)))
    return B


def dsposv(h, A, B, X, upper=False, prec="fp32", tol=1e-14, max_iter=30, stream=None):
    """Mixed-precision SPD solve with FP64 refinement; returns (iters, berr)."""
    n, nrhs = B.shape
    it, be = C.c_int(), C.c_double()
    p = {"fp32": capi.FACTOR_FP32, "bf16x3": capi.FACTOR_BF16X3}[prec]
    capi.check(capi.load().rmh_dsposv(h.value, int(upper), p, n, nrhs, _p(A), _ld(A), _p(B), _ld(B), _p(X), _ld(X),
                                      tol, max_iter, C.byref(it), C.byref(be), _stream(stream)))
    return it.value, be.value


def pstrf(h, A, upper=False, stream=None):
    """In-place pivoted Cholesky; returns (piv, rank) as device int64 tensors
    (piv[j]: the row swapped with j at step j)."""
    n = A.shape[0]
    piv = torch.empty(max(n, 1), dtype=torch.int64, device=A.device)
    rank = torch.empty(1, dtype=torch.int64, device=A.device)
    f = _fn("rmh_", A.dtype, "pstrf")
    capi.check(f(h.value, int(upper), n, _p(A), _ld(A), _p(piv), _p(rank), _stream(stream)))
    return piv[:n], rank


def chol_update(h, L, alpha, beta, v, upper=False, info=None, stream=None):
    """In-place rank-one update of a Cholesky factor: L L^T <- alpha L L^T +
    beta v v^T (v: a 1-D device tensor, any stride).  Returns the device int64
    info tensor (0, or j+1 for the column that would make it indefinite)."""
    n = L.shape[0]
    if info is None:
        info = torch.zeros(1, dtype=torch.int64, device=L.device)
    f = _fn("rmh_", L.dtype, "chol_update")
    inc = v.stride(0) if v.dim() == 1 and v.numel() > 0 else 1
    capi.check(f(h.value, int(upper), n, float(alpha), float(beta), _p(L), _ld(L), _p(v), inc, _p(info),
                 _stream(stream)))
    return info


def potrf_adjoint(h, L, Lbar, Abar=None, upper=False, stream=None):
    """Abar = sym(L^-T Phi(L^T Lbar) L^-1): gradient w.r.t. A of a scalar
    whose gradient w.r.t. the Cholesky factor is Lbar."""
    n = L.shape[0]
    if Abar is None:
        Abar = empty_colmajor(n, n, dtype=L.dtype, device=L.device)
    f = _fn("rmh_", L.dtype, "potrf_adjoint")
    capi.check(f(h.value, int(upper), n, _p(L), _ld(L), _p(Lbar), _ld(Lbar), _p(Abar), _ld(Abar), _stream(stream)))
    return Abar


class Batch:
    """Device-side description of a batch of column-major SPD systems.

    ``As`` and ``Bs`` are lists of column-major device tensors (``colmajor``);
    the pointer, size and leading-dimension arrays are built once on the
    device, so a launch over them can be captured and replayed.
    """

    def __init__(self, As, Bs=None, order=None):
        dev = As[0].device
        self.As, self.Bs = As, Bs
        self.batch = len(As)
        self.dtype = As[0].dtype
        self.max_n = max(int(a.shape[0]) for a in As)
        self.n = torch.tensor([a.shape[0] for a in As], dtype=torch.int64, device=dev)
        self.lda = torch.tensor([max(1, _ld(a)) for a in As], dtype=torch.int64, device=dev)
        self.pa = torch.tensor([a.data_ptr() for a in As], dtype=torch.int64, device=dev)
        self.info = torch.zeros(self.batch, dtype=torch.int64, device=dev)
        self.nrhs = 0
        if Bs is not None:
            self.nrhs = int(Bs[0].shape[1])
            self.ldb = torch.tensor([max(1, _ld(b)) for b in Bs], dtype=torch.int64, device=dev)
            self.pb = torch.tensor([b.data_ptr() for b in Bs], dtype=torch.int64, device=dev)
        self.order = None if order is None else torch.as_tensor(order, dtype=torch.int64).to(dev)

    def _cfg(self, cfg):
        return None if cfg is None else C.byref(cfg)


def _bp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def posv_batched(h, batch: Batch, cfg=None, stream=None):
    f = _fn("rmh_", batch.dtype, "posv_batched")
    capi.check(f(h.value, batch.batch, batch.max_n, _bp(batch.n), _bp(batch.pa), _bp(batch.lda), batch.nrhs,
                 _bp(batch.pb), _bp(batch.ldb), _bp(batch.info), _bp(batch.order), batch._cfg(cfg), _stream(stream)))
    return batch.info


def potrf_batched(h, batch: Batch, cfg=None, stream=None):
    f = _fn("rmh_", batch.dtype, "potrf_batched")
    capi.check(f(h.value, batch.batch, batch.max_n, _bp(batch.n), _bp(batch.pa), _bp(batch.lda), _bp(batch.info),
                 _bp(batch.order), batch._cfg(cfg), _stream(stream)))
    return batch.info


def potrs_batched(h, batch: Batch, cfg=None, stream=None):
    f = _fn("rmh_", batch.dtype, "potrs_batched")
    capi.check(f(h.value, batch.batch, batch.max_n, _bp(batch.n), _bp(batch.pa), _bp(batch.lda), batch.nrhs,
                 _bp(batch.pb), _bp(batch.ldb), _bp(batch.order), batch._cfg(cfg), _stream(stream)))


def lpt_order(sizes):
    import numpy as np
    n = np.ascontiguousarray(np.asarray(sizes, dtype=np.int64))
    out = np.empty_like(n)
    capi.check(capi.load().rmh_batched_lpt_order(len(n), n.ctypes.data, out.ctypes.data))
    return out


def batched_autotune(h, sizes, nrhs, fp64=True):
    """Run the autotuner on host sizes; returns (BatchConfig, best ms)."""
    import numpy as np
    n = np.ascontiguousarray(np.asarray(sizes, dtype=np.int64))
    cfg = capi.BatchConfig()
    ms = C.c_double()
    capi.check(capi.load().rmh_batched_autotune(h.value, int(fp64), len(n), n.ctypes.data, nrhs, C.byref(cfg),
                                                C.byref(ms)))
    return cfg, ms.value


def set_potrf_blocking(h, nb=0, nb_inner=0):
    capi.check(capi.load().rmh_set_potrf_blocking(h.value, nb, nb_inner))


def set_potrf_leaf(h, leaf=0):
    """Leaf width of the recursive panel factorisation (0: the default)."""
    capi.check(capi.load().rmh_set_potrf_leaf(h.value, leaf))


def set_potrf_reserve(h, cus_per_xcd=0, rows_lo=0, rows_hi=1 << 62):
    """Compute units per XCD the look-ahead trailing updates leave to the
    panel launches (0: none, the default), for the panel steps whose
    trailing rows n - k1 lie in (rows_lo, rows_hi]."""
    capi.check(capi.load().rmh_set_potrf_reserve(h.value, cus_per_xcd, rows_lo, rows_hi))


def set_spotrf_precision(h, prec="fp32"):
    p = {"fp32": capi.FACTOR_FP32, "bf16x3": capi.FACTOR_BF16X3}[prec]
    capi.check(capi.load().rmh_set_spotrf_precision(h.value, p))


def set_potrf_lookahead(h, enable=-1):
    """enable: False / True, or the mode 0 (off), 1 (on), 2 (on, trailing
    update one workgroup per compute unit), 3 / 4 (modes 2 / 1 with
    four-wavefront panel workgroups); -1 (the default) picks by precision."""
    capi.check(capi.load().rmh_set_potrf_lookahead(h.value, int(enable)))


class Profile:
    """Launch profiler of a handle (HIP events around each launch class)."""

    CLASSES = ("trailing_update", "panel", "trailing_update_small")

    def __init__(self, h):
        self.h = h
        self.lib = capi.load()

    def enable(self, classes=CLASSES):
        mask = 0
        for c in classes:
            mask |= 1 << self.CLASSES.index(c)
        capi.check(self.lib.rmh_profile_enable(self.h.value, mask))

    def disable(self):
        capi.check(self.lib.rmh_profile_enable(self.h.value, 0))

    def reset(self):
        capi.check(self.lib.rmh_profile_reset(self.h.value))

    def read(self, cls: str):
        ms, fl, n = C.c_double(), C.c_double(), C.c_int64()
        capi.check(self.lib.rmh_profile_read(self.h.value, cls.encode(), C.byref(ms), C.byref(fl), C.byref(n)))
        return {"ms": ms.value, "flops": fl.value, "launches": n.value}
