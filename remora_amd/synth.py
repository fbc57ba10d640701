"""Seeded synthetic inputs for the Cholesky workloads (host side, numpy).

Random numbers come from SplitMix64, which is counter-based: element i of
stream s under seed k is mix(key(k, s) + (i+1)*GAMMA), so a slice of a stream
can be generated anywhere (the GPU generators in the library use the same
integer function).  Uniforms take the top 53 bits; normals use Box-Muller on
consecutive uniform pairs.  std::normal_distribution is not used because it
is implementation-defined.
"""
from __future__ import annotations

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_STREAM = np.uint64(0xD1B54A32D192ED03)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def stream_key(seed: int, stream: int = 0) -> np.uint64:
    with np.errstate(over="ignore"):
        k = np.array([np.uint64(seed) ^ (np.uint64(stream) * _STREAM)], dtype=np.uint64)
        return _mix(k + GAMMA)[0]


def raw(seed: int, stream: int, count: int, offset: int = 0) -> np.ndarray:
    key = stream_key(seed, stream)
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        return _mix(key + idx * GAMMA)


def uniform(seed: int, stream: int, count: int, lo=0.0, hi=1.0, offset: int = 0) -> np.ndarray:
    u = (raw(seed, stream, count, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def normal(seed: int, stream: int, count: int) -> np.ndarray:
    m = (count + 1) // 2
    u = uniform(seed, stream, 2 * m)
    u1, u2 = 1.0 - u[0::2], u[1::2]  # u1 in (0, 1]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2.0 * np.pi * u2)
    z[1::2] = r * np.sin(2.0 * np.pi * u2)
    return z[:count]


def gram_spd(n: int, seed: int = 0, divisor: float | None = None, shift: float = 0.1) -> np.ndarray:
    """A = G G^T / divisor + shift I with G n x n N(0,1) (row-major fill).

    divisor defaults to n (spectrum O(1)); the 512-point CPU config uses 512."""
    g = normal(seed, 0, n * n).reshape(n, n)
    a = g @ g.T / float(divisor if divisor is not None else n)
    a[np.diag_indices(n)] += shift
    return a


def rhs(n: int, nrhs: int, seed: int = 0, stream: int = 1) -> np.ndarray:
    """B n x nrhs, U[-1, 1] (row-major fill)."""
    return uniform(seed, stream, n * nrhs, -1.0, 1.0).reshape(n, nrhs)


def points(n: int, d: int, seed: int, stream: int = 0) -> np.ndarray:
    return uniform(seed, stream, n * d).reshape(n, d)


def rbf(x: np.ndarray, lengthscale: float, jitter: float) -> np.ndarray:
    """A_ij = exp(-|x_i - x_j|^2 / (2 l^2)) + jitter * delta_ij."""
    sq = np.sum(x * x, axis=1)
    d2 = np.maximum(sq[:, None] + sq[None, :] - 2.0 * (x @ x.T), 0.0)
    a = np.exp(-d2 / (2.0 * lengthscale * lengthscale))
    a[np.diag_indices(len(x))] += jitter
    return a


def batched_sizes(count: int = 4096, lo: int = 16, hi: int = 512, seed: int = 2) -> np.ndarray:
    """n_i uniform-integer in [lo, hi]."""
    u = uniform(seed, 0, count)
    return (lo + np.floor(u * (hi - lo + 1))).astype(np.int64)


def backward_error(A: np.ndarray, X: np.ndarray, B: np.ndarray) -> float:
    """max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)."""
    R = B - A @ X
    an = np.max(np.sum(np.abs(A), axis=1))
    num = np.max(np.abs(R), axis=0)
    den = an * np.max(np.abs(X), axis=0) + np.max(np.abs(B), axis=0)
    return float(np.max(num / den))


def factor_error(A: np.ndarray, L: np.ndarray) -> float:
    """||A - L L^T||_F / ||A||_F with L the lower triangle of the argument."""
    Lt = np.tril(L)
    return float(np.linalg.norm(A - Lt @ Lt.T) / np.linalg.norm(A))
