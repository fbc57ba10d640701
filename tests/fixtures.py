"""Closed-form test matrices of the reference's own tests (restated).

Each generator cites the reference test it follows; they are deterministic
closed forms, not random draws.
"""
from __future__ import annotations

import numpy as np


def create_symm(n: int, rank: int = 0, hip_variant: bool = False) -> np.ndarray:
    """createSymm: A = R R^T, R(i,j) = 0.2/|i-j| below the diagonal,
    R(i,i) = 0.5/n*i + 1 (Test/cholesky.cpp:15-34, Test/symm_solve.cpp:17-36)
    or 1 + 0.5/n*i + 1 (Test/hip_cholesky.cpp:14-33)."""
    if rank == 0:
        rank = n
    R = np.zeros((n, n))
    for i in range(n):
        for j in range(min(i, rank)):
            R[i, j] = 0.2 / abs(i - j)
        if i < rank:
            R[i, i] = (1.0 if hip_variant else 0.0) + 0.5 / n * i + 1
    A = R @ R.T
    if rank != n:
        for i in range(rank // 2):
            a, b = 2 * i, n - i - 1
            A[[a, b], :] = A[[b, a], :]
            A[:, [a, b]] = A[:, [b, a]]
    return A


def syrk_arg(dims: int = 384, K: int = 244) -> np.ndarray:
    """arg(i,j) = i/dims + 0.2/K*j + 1 (Test/syrk.cpp:53-59)."""
    i = np.arange(dims)[:, None]
    j = np.arange(K)[None, :]
    return (1.0 / dims) * i + 0.2 / K * j + 1.0


def triangular_matrix(size: int = 158, matrix_case: bool = True) -> np.ndarray:
    """Lower-triangular A with ones above the diagonal as filler
    (Test/triangular_solve.cpp:22-29 vector case, :120-127 matrix case)."""
    A = np.ones((size, size))
    for i in range(size):
        for j in range(i + 1):
            if matrix_case:
                A[i, j] = 0.2 / size * i - 0.05 / (i + 1.0) * j + (10.0 if i == j else 0.0)
            else:
                A[i, j] = 0.1 / size * i - 0.05 / (i + 1.0) * j + (0.5 if i == j else 0.0)
    return A


def triangular_rhs(size: int = 158, k: int = 138) -> np.ndarray:
    """B(i,j) = 0.1/size*i + 0.1*k (Test/triangular_solve.cpp:129-134)."""
    i = np.arange(size)[:, None]
    return np.broadcast_to((0.1 / size) * i + 0.1 * k, (size, k)).copy()


def symm_solve_rhs(n: int = 128, k: int = 151) -> np.ndarray:
    """B(i,j) = 0.1*j + i/n (Test/symm_solve.cpp:95-100)."""
    i = np.arange(n)[:, None]
    j = np.arange(k)[None, :]
    return 0.1 * j + (1.0 / n) * i


def prod_operands(rows=80, middle=33, columns=50):
    """arg1(i,j) = i*middle + 0.2*j, arg2(i,j) = i*columns + 1.5*j
    (Test/prod.cpp:353-367); the cases compute result(1.5) += -2 * arg1 arg2."""
    i = np.arange(rows)[:, None]
    j = np.arange(middle)[None, :]
    A = i * middle + 0.2 * j
    i2 = np.arange(middle)[:, None]
    j2 = np.arange(columns)[None, :]
    B = i2 * columns + 1.5 * j2
    return A.astype(np.float64), B.astype(np.float64)


def triangular_vector_rhs(size: int = 158) -> np.ndarray:
    """b(i) = 0.1/size*i (Test/triangular_solve.cpp:31-34)."""
    return (0.1 / size) * np.arange(size, dtype=np.float64)


def prod_vector_operands(rows: int = 50, columns: int = 40):
    """arg1(i,j) = 0.01*i*columns + 0.1*j, arg2(j) = 0.1*j + 0.1
    (Test/hip_prod.cpp:33-49); the cases compute result(1.5) += -2 * arg1 arg2."""
    i = np.arange(rows)[:, None]
    j = np.arange(columns)[None, :]
    A = 0.01 * i * columns + 0.1 * j
    x = 0.1 * np.arange(columns) + 0.1
    return A.astype(np.float64), x.astype(np.float64)
