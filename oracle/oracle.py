"""ctypes front of the CPU oracle (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, and only as the checker: the HIP library never
calls it.  Each function restates the Remora CPU entry point named in its
docstring (paths relative to the Remora tree) on numpy arrays; the arithmetic
runs in ``oracle/remora_cpu.c``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def load(path: str | None = None):
    """Load (building if absent) the oracle shared library."""
    global _lib
    if path is not None:
        lib = C.CDLL(path)
        _declare(lib)
        return lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from oracle import build as _b
            _b.build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(lib):
    P, S, I, D = C.c_void_p, C.c_size_t, C.c_int, C.c_double
    lib.orc_dpotrf.argtypes = [P, S, S, I, I]
    lib.orc_dpotrf.restype = S
    lib.orc_dtrsm.argtypes = [P, S, S, I, P, S, S, S, I, I, I, I]
    lib.orc_dtrsm.restype = I
    lib.orc_dsyrk.argtypes = [P, S, S, S, I, P, S, I, D, I]
    lib.orc_dsyrk.restype = None
    lib.orc_dgemm.argtypes = [P, S, I, P, S, I, P, S, I, S, S, S, D]
    lib.orc_dgemm.restype = None
    lib.orc_dposv.argtypes = [P, S, S, I, P, S, S, S, I, I, C.POINTER(S)]
    lib.orc_dposv.restype = I
    lib.orc_dchol_update.argtypes = [P, S, S, I, D, D, P, S]
    lib.orc_dchol_update.restype = S
    lib.orc_dpstrf.argtypes = [P, S, S, I, I, P]
    lib.orc_dpstrf.restype = S
    Q = C.c_ssize_t
    lib.orc_dtrsv.argtypes = [P, S, S, I, P, Q, I, I, I]
    lib.orc_dtrsv.restype = I
    lib.orc_dgemv.argtypes = [P, S, S, S, I, P, Q, P, Q, D]
    lib.orc_dgemv.restype = None
    lib.orc_dposv_vec.argtypes = [P, S, S, I, P, Q, C.POINTER(S)]
    lib.orc_dposv_vec.restype = I
    lib.orc_vector_length.restype = I
    lib.orc_simd_enabled.restype = I


def _prep(a: np.ndarray, row_major: bool) -> np.ndarray:
    """Copy to a float64 array in the requested memory order."""
    return np.array(a, dtype=np.float64, order="C" if row_major else "F", copy=True)


def _ld(a: np.ndarray, row_major: bool) -> int:
    return a.strides[0] // 8 if row_major else a.strides[1] // 8


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def potrf(A, upper: bool = False, row_major: bool = True, lib=None):
    """kernels::potrf<lower|upper>(A) (kernels/default/potrf.hpp:148-155).

    Returns (factored copy of A, info).  Only the requested triangle is
    overwritten; info is Remora's leaf-relative failing pivot (0 = success).
    """
    lib = lib or load()
    a = _prep(A, row_major)
    n = a.shape[0]
    info = lib.orc_dpotrf(_ptr(a), n, _ld(a, row_major) if n else 1, int(row_major), int(upper))
    return a, int(info)


def trsm(A, B, upper: bool, unit: bool, left: bool, a_row_major=True, b_row_major=True, lib=None):
    """kernels::trsm<Triangular,Side>(A, B) (kernels/default/trsm.hpp:278-286).

    Solves op-free A X = B (left) or X A = B (right) in place on a copy of B.
    Returns (X, singular) where singular mirrors the reference's
    std::invalid_argument on a zero diagonal.
    """
    lib = lib or load()
    a = _prep(A, a_row_major)
    b = _prep(B, b_row_major)
    n = a.shape[0]
    sing = lib.orc_dtrsm(_ptr(a), n, max(_ld(a, a_row_major), 1), int(a_row_major), _ptr(b), b.shape[0],
                         b.shape[1], max(_ld(b, b_row_major), 1), int(b_row_major), int(upper), int(unit),
                         int(left))
    return b, bool(sing)


def syrk(E, M, alpha: float, upper: bool, e_row_major=True, m_row_major=True, lib=None):
    """kernels::syrk<Upper>(E, M, alpha): M += alpha E E^T on one triangle
    (kernels/default/syrk.hpp:50-130).  Returns the updated copy of M."""
    lib = lib or load()
    e = _prep(E, e_row_major)
    m = _prep(M, m_row_major)
    lib.orc_dsyrk(_ptr(e), e.shape[0], e.shape[1], max(_ld(e, e_row_major), 1), int(e_row_major), _ptr(m),
                  max(_ld(m, m_row_major), 1), int(m_row_major), float(alpha), int(upper))
    return m


def gemm(E1, E2, M, alpha: float, e1_row_major=True, e2_row_major=True, m_row_major=True, lib=None):
    """kernels::gemm(E1, E2, M, alpha): M += alpha E1 E2
    (kernels/gemm.hpp:85-106 -> default/dense_gemm.hpp:83-141)."""
    lib = lib or load()
    e1 = _prep(E1, e1_row_major)
    e2 = _prep(E2, e2_row_major)
    m = _prep(M, m_row_major)
    mm, kk = e1.shape
    nn = e2.shape[1]
    lib.orc_dgemm(_ptr(e1), max(_ld(e1, e1_row_major), 1), int(e1_row_major), _ptr(e2),
                  max(_ld(e2, e2_row_major), 1), int(e2_row_major), _ptr(m), max(_ld(m, m_row_major), 1),
                  int(m_row_major), mm, nn, kk, float(alpha))
    return m


def posv(A, B, left: bool = True, a_row_major=True, b_row_major=True, lib=None):
    """solve(A, B, symm_pos_def(), left|right) through cholesky_decomposition
    (solve.hpp:165-170, decompositions.hpp:81-110).

    Returns (X, L_with_upper_of_A, info, singular)."""
    lib = lib or load()
    a = _prep(A, a_row_major)
    b = _prep(B, b_row_major)
    n = a.shape[0]
    info = C.c_size_t(0)
    sing = lib.orc_dposv(_ptr(a), n, max(_ld(a, a_row_major), 1), int(a_row_major), _ptr(b), b.shape[0],
                         b.shape[1], max(_ld(b, b_row_major), 1), int(b_row_major), int(left), C.byref(info))
    return b, a, int(info.value), bool(sing)


def chol_update(L, alpha: float, beta: float, v, row_major: bool = True, incv: int = 1, lib=None):
    """cholesky_decomposition::update(alpha, beta, v) on a lower factor
    (decompositions.hpp:132-169).  v holds n elements at stride incv.

    Returns (updated copy of L, info): info 0, or j+1 when column j would make
    the matrix indefinite (the reference throws there; the copy then holds
    columns 0..j-1 updated, as the reference's matrix does)."""
    lib = lib or load()
    a = _prep(L, row_major)
    n = a.shape[0]
    vv = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
    assert vv.size >= (n - 1) * incv + 1 or n == 0
    info = lib.orc_dchol_update(_ptr(a), n, max(_ld(a, row_major), 1) if n else 1, int(row_major), float(alpha),
                                float(beta), _ptr(vv), incv)
    return a, int(info)


def pstrf(A, upper: bool = False, row_major: bool = True, lib=None):
    """kernels::pstrf<lower|upper>(A, P) (kernels/default/pstrf.hpp:40-160).

    Returns (factored copy of A, P, rank): P starts as the identity and holds
    the row swapped with j at step j (the reference's permutation_matrix)."""
    lib = lib or load()
    a = _prep(A, row_major)
    n = a.shape[0]
    P = np.arange(n, dtype=np.int64)
    rank = lib.orc_dpstrf(_ptr(a), n, max(_ld(a, row_major), 1) if n else 1, int(row_major), int(upper),
                          _ptr(P) if n else None)
    return a, P, int(rank)


def swap_full_inverted(P, M):
    """P^T M P for a transposition record P (the reference's
    swap_full_inverted: swaps applied in reverse order to rows and columns)."""
    M = np.array(M, copy=True)
    for i in range(len(P) - 1, -1, -1):
        p = int(P[i])
        if p != i:
            M[[i, p], :] = M[[p, i], :]
            M[:, [i, p]] = M[:, [p, i]]
    return M


def _strided(v, inc: int) -> np.ndarray:
    """A float64 buffer holding v at stride inc (gaps filled with NaN)."""
    v = np.asarray(v, dtype=np.float64).ravel()
    buf = np.full(max(1, (len(v) - 1) * inc + 1), np.nan)
    buf[::inc] = v
    return buf


def trsv(A, b, upper: bool, unit: bool, left: bool = True, a_row_major=True, inc: int = 1, lib=None):
    """kernels::trsv<Triangular, Side>(A, b) (kernels/trsv.hpp:53-63 ->
    default/trsv.hpp:44-183) on a copy of b stored at stride inc.

    Returns (x, singular): singular mirrors the reference's
    std::invalid_argument on a zero diagonal."""
    lib = lib or load()
    a = _prep(A, a_row_major)
    n = a.shape[0]
    buf = _strided(b, inc)
    sing = lib.orc_dtrsv(_ptr(a), n, max(_ld(a, a_row_major), 1) if n else 1, int(a_row_major), _ptr(buf), inc,
                         int(upper), int(unit), int(left))
    return buf[::inc][:n].copy(), bool(sing)


def gemv(A, x, y, alpha: float, a_row_major=True, lib=None):
    """kernels::gemv(A, x, y, alpha): y += alpha A x (kernels/gemv.hpp:58-72
    -> default/gemv.hpp:43-93).  Returns the updated copy of y."""
    lib = lib or load()
    a = _prep(A, a_row_major)
    m, n = a.shape
    xx = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    yy = np.array(y, dtype=np.float64, copy=True)
    lib.orc_dgemv(_ptr(a), m, n, max(_ld(a, a_row_major), 1) if m and n else 1, int(a_row_major), _ptr(xx), 1,
                  _ptr(yy), 1, float(alpha))
    return yy


def posv_vec(A, b, a_row_major=True, lib=None):
    """solve(A, b, symm_pos_def(), left|right) for a vector b through
    cholesky_decomposition (solve.hpp:90-101, decompositions.hpp:81-91,
    111-115).  Returns (x, L_with_upper_of_A, info, singular)."""
    lib = lib or load()
    a = _prep(A, a_row_major)
    n = a.shape[0]
    bb = np.array(b, dtype=np.float64, copy=True).ravel()
    info = C.c_size_t(0)
    sing = lib.orc_dposv_vec(_ptr(a), n, max(_ld(a, a_row_major), 1) if n else 1, int(a_row_major), _ptr(bb), 1,
                             C.byref(info))
    return bb, a, int(info.value), bool(sing)


def vector_length(lib=None) -> int:
    return int((lib or load()).orc_vector_length())


def simd_enabled(lib=None) -> bool:
    return bool((lib or load()).orc_simd_enabled())
