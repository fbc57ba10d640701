"""TEST INFRASTRUCTURE ONLY: CPU FP64 reference of the Cholesky adjoint.

The reference library has no derivative of its factorisation; this is the
textbook reverse-mode rule (I. Murray, "Differentiation of the Cholesky
decomposition", 2016, eq. 10 with the symmetric-input convention): for
A = L L^T and the gradient Lbar with respect to the lower-triangular L,

    P = Phi(L^T Lbar),  Phi = lower triangle with the diagonal halved,
    G = L^-T P L^-1,    Abar = (G + G^T) / 2.

Only tests call it, as the checker of rmh_?potrf_adjoint; it is pinned in
tests/test_adjoint_ref.py against torch's own Cholesky backward, central
finite differences and the closed-form gradient of a Gaussian-process
negative log marginal likelihood.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import solve_triangular


def phi(X: np.ndarray) -> np.ndarray:
    P = np.tril(X)
    P[np.diag_indices(len(P))] *= 0.5
    return P


def potrf_adjoint(L: np.ndarray, Lbar: np.ndarray) -> np.ndarray:
    """Symmetric gradient w.r.t. A of a scalar with gradient Lbar w.r.t. L."""
    L = np.tril(L)
    P = phi(L.T @ np.tril(Lbar))
    X = solve_triangular(L.T, P, lower=False)            # L^-T P
    G = solve_triangular(L.T, X.T, lower=False).T        # X L^-1
    return 0.5 * (G + G.T)


def gp_nlml(A: np.ndarray, y: np.ndarray) -> float:
    """0.5 y^T A^-1 y + sum log diag(L)."""
    L = np.linalg.cholesky(A)
    z = solve_triangular(L, y, lower=True)
    return float(0.5 * z @ z + np.sum(np.log(np.diag(L))))


def gp_nlml_grad(A: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Closed form: 0.5 (A^-1 - alpha alpha^T), alpha = A^-1 y."""
    Ainv = np.linalg.inv(A)
    alpha = Ainv @ y
    return 0.5 * (Ainv - np.outer(alpha, alpha))


def gp_nlml_lbar(L: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Gradient of the NLML w.r.t. the lower-triangular factor L:
    f = 0.5 |L^-1 y|^2 + sum log L_ii gives Lbar = tril(-L^-T z z^T L^-T ... )
    written as Lbar = tril(-w z^T) + diag(1/L_ii), z = L^-1 y, w = L^-T z."""
    z = solve_triangular(L, y, lower=True)
    w = solve_triangular(L.T, z, lower=False)
    Lbar = np.tril(-np.outer(w, z))
    Lbar[np.diag_indices(len(L))] += 1.0 / np.diag(L)
    return Lbar
