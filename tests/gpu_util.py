"""Host <-> device helpers for the GPU tests (column-major CUDA tensors)."""
from __future__ import annotations

import numpy as np


def dev(a: np.ndarray, dtype=None, ld: int | None = None):
    """Upload a 2-D numpy array as a column-major CUDA tensor.  With ld > rows
    the matrix is a view into a taller buffer (a padded leading dimension)."""
    import torch
    a = np.asarray(a)
    if dtype is None:
        dtype = torch.float64 if a.dtype == np.float64 else torch.float32
    rows, cols = a.shape
    ld = rows if ld is None else ld
    buf = torch.full((max(cols, 1), max(ld, 1)), float("nan"), dtype=dtype)
    buf[:cols, :rows] = torch.from_numpy(np.ascontiguousarray(a.T)).to(dtype)
    return buf.cuda().t()[:rows, :cols]


def host(t) -> np.ndarray:
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().to(torch.float64).numpy().copy()


def componentwise_ok(got, ref, scale, tol):
    """|got - ref| <= tol * scale elementwise (scale: the |A||B|-type bound)."""
    err = np.abs(got - ref)
    bound = tol * np.maximum(scale, np.finfo(np.float64).tiny)
    worst = float(np.max(err / bound)) if err.size else 0.0
    return bool(np.all(err <= bound)), worst
