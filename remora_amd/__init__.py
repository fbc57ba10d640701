"""remora_amd -- MI355X (gfx950) implementation of Remora's dense Cholesky path.

The product is the C-ABI library ``remora_amd/lib/libremora_hip.so``
(``include/remora_hip/capi.h``) and the header-only C++ front-end under
``include/remora/``.  This package holds the HIP sources (``csrc/``), the
build recipe and a thin Python binding used by the tests and the benchmark.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
