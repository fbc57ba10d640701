"""Pins of the CPU oracle (oracle/remora_cpu.c), no GPU needed.

The reference ships no golden vectors for this path; its tests are closed-form
generators plus invariants.  The oracle is pinned by
  * those reference tests, restated (fixtures cited per test);
  * LAPACK (scipy) and 50-digit mpmath Cholesky / triangular solves;
  * special cases (identity, diagonal, failing pivots);
  * planted mistakes: each mutant build of the oracle must fail a pin.
"""
import os

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import build as orc_build
from oracle import oracle as orc
from remora_amd import synth
from tests import fixtures

ORIENT = [True, False]


def _mp_cholesky(A, dps=50):
    import mpmath
    mpmath.mp.dps = dps
    n = A.shape[0]
    M = mpmath.matrix(A.tolist())
    L = mpmath.matrix(n, n)
    for j in range(n):
        s = M[j, j] - sum(L[j, k] ** 2 for k in range(j))
        L[j, j] = mpmath.sqrt(s)
        for i in range(j + 1, n):
            L[i, j] = (M[i, j] - sum(L[i, k] * L[j, k] for k in range(j))) / L[j, j]
    return np.array([[float(L[i, j]) for j in range(n)] for i in range(n)])


# ---------------------------------------------------------------- potrf --
@pytest.mark.parametrize("row_major", ORIENT)
def test_potrf_reference_fixture(row_major):
    """Test/cholesky.cpp:40-80 on createSymm(123)."""
    A = fixtures.create_symm(123)
    low, i1 = orc.potrf(A, upper=False, row_major=row_major)
    up, i2 = orc.potrf(A, upper=True, row_major=row_major)
    assert i1 == 0 and i2 == 0
    iu = np.triu_indices(123, 1)
    np.testing.assert_allclose(low[iu], A[iu], rtol=1e-14)
    np.testing.assert_allclose(up.T[iu], A.T[iu], rtol=1e-14)
    L, U = np.tril(low), np.triu(up)
    assert np.max(np.abs(A - L @ L.T)) < 1e-12
    assert np.max(np.abs(A - U.T @ U)) < 1e-12
    assert np.all(np.isfinite(L)) and np.all(np.isfinite(U))
    # cholesky_decomposition yields the same factor (decompositions.hpp:81-84)
    _, Ldec, info, _ = orc.posv(A, np.zeros((123, 1)), a_row_major=row_major)
    assert info == 0
    np.testing.assert_allclose(np.tril(Ldec), L, rtol=1e-14)
    np.testing.assert_allclose(U, L.T, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("n", [1, 2, 31, 32, 33, 64, 123, 200, 517])
@pytest.mark.parametrize("row_major", ORIENT)
def test_potrf_vs_lapack(n, row_major):
    A = synth.gram_spd(n, seed=n, divisor=max(n, 1))
    F, info = orc.potrf(A, row_major=row_major)
    assert info == 0
    L = np.tril(F)
    Llap = sla.cholesky(A, lower=True)
    scale = np.sqrt(np.outer(np.diag(A), np.diag(A)))
    assert np.max(np.abs(L - Llap) / scale) < 50 * n * np.finfo(float).eps


@pytest.mark.parametrize("n", [5, 33, 40])
def test_potrf_vs_mpmath(n):
    """Exact (50-digit) Cholesky: the oracle is within c*n*u of it."""
    A = fixtures.create_symm(n)
    F, info = orc.potrf(A)
    assert info == 0
    Lx = _mp_cholesky(A)
    scale = np.sqrt(np.outer(np.diag(A), np.diag(A)))
    assert np.max(np.abs(np.tril(F) - Lx) / scale) < 4 * n * np.finfo(float).eps


def test_potrf_special_cases():
    I = np.eye(70)
    F, info = orc.potrf(I)
    assert info == 0 and np.array_equal(np.tril(F), I)
    d = np.arange(1.0, 71.0)
    F, info = orc.potrf(np.diag(d), upper=True)
    assert info == 0
    np.testing.assert_allclose(np.diag(F), np.sqrt(d), rtol=0, atol=0)
    _, info = orc.potrf(-np.eye(10))
    assert info == 1


def test_potrf_info_is_leaf_relative():
    """The reference's info is the failing row inside its 32-wide leaf
    (default/potrf.hpp:55-56 on the leaf view; the recursion returns it
    unchanged).  A failure at global row 40 (leaf 32..63) reports 9."""
    n = 100
    A = synth.gram_spd(n, seed=9)
    L = np.linalg.cholesky(A[:41, :41])
    A[40, 40] -= 1.5 * L[40, 40] ** 2
    _, info = orc.potrf(A)
    assert info == 40 - 32 + 1
    _, info = orc.potrf(A, row_major=False)
    assert info != 0


# ----------------------------------------------------------------- trsm --
TRSM_CASES = [(left, upper, unit) for left in (True, False) for upper in (False, True) for unit in (False, True)]


@pytest.mark.parametrize("left,upper,unit", TRSM_CASES)
@pytest.mark.parametrize("b_row_major", ORIENT)
def test_trsm_reference_fixture(left, upper, unit, b_row_major):
    """Test/triangular_solve.cpp:117-212: residual < 1e-11 and agreement
    with an independent triangular solver (LAPACK trtrs via scipy)."""
    Alow = fixtures.triangular_matrix(158, True)
    A = Alow if not upper else Alow.T.copy()
    B = fixtures.triangular_rhs(158, 138)
    if not left:
        B = B.T.copy()
    X, sing = orc.trsm(A, B, upper, unit, left, b_row_major=b_row_major)
    assert not sing
    T = np.triu(A) if upper else np.tril(A)
    if unit:
        T = T - np.diag(np.diag(T)) + np.eye(158)
    R = T @ X - B if left else X @ T - B
    assert np.max(np.abs(R)) < 1e-11
    if left:
        ref = sla.solve_triangular(T, B, lower=not upper)
    else:
        ref = sla.solve_triangular(T.T, B.T, lower=upper).T
    np.testing.assert_allclose(X, ref, rtol=1e-12, atol=1e-14 * np.max(np.abs(ref)))


def test_trsm_vs_mpmath_small():
    import mpmath
    mpmath.mp.dps = 40
    n = 37
    A = np.tril(synth.uniform(1, 0, n * n, -1, 1).reshape(n, n)) + 3 * np.eye(n)
    b = synth.uniform(1, 1, n, -1, 1).reshape(n, 1)
    X, _ = orc.trsm(A, b, False, False, True)
    M = mpmath.matrix(A.tolist())
    x = mpmath.lu_solve(M, mpmath.matrix(b.tolist()))
    xe = np.array([float(v) for v in x]).reshape(n, 1)
    assert np.max(np.abs(X - xe)) < 1e-14 * np.max(np.abs(xe))


def test_trsm_singular_is_reported():
    """A zero diagonal makes the reference throw std::invalid_argument
    (default/trsm.hpp:86-87)."""
    A = np.tril(np.ones((50, 50)))
    A[20, 20] = 0.0
    _, sing = orc.trsm(A, np.ones((50, 3)), False, False, True)
    assert sing
    _, sing = orc.trsm(A, np.ones((50, 3)), False, True, True)
    assert not sing


# ----------------------------------------------------------------- syrk --
@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("e_rm", ORIENT)
@pytest.mark.parametrize("m_rm", ORIENT)
def test_syrk_reference_fixture(upper, e_rm, m_rm):
    """Test/syrk.cpp:13-87: closed form alpha<row_i,row_j> + init on the
    touched triangle at 1e-10 percent, the other triangle untouched."""
    arg = fixtures.syrk_arg()
    dims = arg.shape[0]
    M = orc.syrk(arg, np.full((dims, dims), 3.0), 2.0, upper, e_row_major=e_rm, m_row_major=m_rm)
    full = 2.0 * arg @ arg.T + 3.0
    mask = np.triu(np.ones((dims, dims), bool)) if upper else np.tril(np.ones((dims, dims), bool))
    np.testing.assert_allclose(M[mask], full[mask], rtol=1e-12)
    assert np.all(M[~mask] == 3.0)


# ----------------------------------------------------------------- gemm --
@pytest.mark.parametrize("e1_rm", ORIENT)
@pytest.mark.parametrize("e2_rm", ORIENT)
@pytest.mark.parametrize("m_rm", ORIENT)
def test_gemm_reference_fixture(e1_rm, e2_rm, m_rm):
    """Test/prod.cpp:330-472: result(1.5) += -2 * arg1 * arg2 for all eight
    orientation combinations against the textbook triple loop."""
    A, B = fixtures.prod_operands()
    got = orc.gemm(A, B, np.full((80, 50), 1.5), -2.0, e1_rm, e2_rm, m_rm)
    ref = np.full((80, 50), 1.5)
    for i in range(80):
        for j in range(50):
            s = 1.5
            for k in range(33):
                s += -2.0 * A[i, k] * B[k, j]
            ref[i, j] = s
    np.testing.assert_allclose(got, ref, rtol=1e-12)


@pytest.mark.parametrize("mnk", [(1, 1, 1), (257, 1031, 600), (129, 13, 1100)])
def test_gemm_blocking_edges(mnk):
    """Sizes that cross the mc=128 / kc=512 / nc=1020 blocks."""
    m, n, k = mnk
    A = synth.normal(2, 0, m * k).reshape(m, k)
    B = synth.normal(2, 1, k * n).reshape(k, n)
    got = orc.gemm(A, B, np.zeros((m, n)), 1.0)
    scale = np.abs(A) @ np.abs(B)
    assert np.all(np.abs(got - A @ B) <= 4 * k * np.finfo(float).eps * scale)


# ---------------------------------------------------------------- solve --
@pytest.mark.parametrize("left", [True, False])
@pytest.mark.parametrize("b_rm", ORIENT)
def test_symm_solve_reference_fixture(left, b_rm):
    """Test/symm_solve.cpp:89-139 (n=128, k=151): max|A X - B| < 1e-12."""
    A = fixtures.create_symm(128)
    B = fixtures.symm_solve_rhs(128, 151)
    if not left:
        B = B.T.copy()
    X, _, info, sing = orc.posv(A, B, left=left, b_row_major=b_rm)
    assert info == 0 and not sing
    R = A @ X - B if left else X @ A - B
    assert np.max(np.abs(R)) < 1e-12
    ref = np.linalg.solve(A, B) if left else np.linalg.solve(A.T, B.T).T
    np.testing.assert_allclose(X, ref, rtol=1e-12, atol=1e-14)


def test_cpu_config_backward_error():
    """The 512-point CPU configuration: A = G G^T/512 + 0.1 I (seed 0),
    16 RHS U[-1,1]."""
    A = synth.gram_spd(512, seed=0, divisor=512)
    B = synth.rhs(512, 16, seed=0)
    X, L, info, _ = orc.posv(A, B, a_row_major=True, b_row_major=True)
    assert info == 0
    assert synth.factor_error(A, L) < 1e-15
    assert synth.backward_error(A, X, B) < 1e-16


def test_block_constants():
    """AVX build: 32-byte vectors (simd.hpp:49-53) with REMORA_USE_SIMD."""
    assert orc.vector_length() == 32
    assert orc.simd_enabled()


# --------------------------------------------------------------- update --
def _update_case(n, alpha, beta, seed=0):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n))
    A = G @ G.T / n + 0.5 * np.eye(n)
    u = rng.standard_normal(n)
    return A, 2.0 * u / np.linalg.norm(u)


@pytest.mark.parametrize("row_major", ORIENT)
def test_update_reference_fixture(row_major):
    """Test/cholesky.cpp:146-167: n = 123, v_i = 0.1/123*i - 0.5, alpha 0.1,
    beta 0.5; the updated factor equals the factor of alpha A + beta v v^T to
    1e-11 percent."""
    n = 123
    A = fixtures.create_symm(n)
    v = 0.1 / n * np.arange(n) - 0.5
    F, info = orc.potrf(A, row_major=row_major)
    assert info == 0
    U, uinfo = orc.chol_update(F, 0.1, 0.5, v, row_major=row_major)
    R, rinfo = orc.potrf(0.1 * A + 0.5 * np.outer(v, v), row_major=row_major)
    assert uinfo == 0 and rinfo == 0
    lo = np.tril_indices(n)
    assert np.allclose(U[lo], R[lo], rtol=1e-13, atol=0)
    # the strict upper triangle (A's values in the reference's storage) is untouched
    up = np.triu_indices(n, 1)
    assert np.array_equal(U[up], F[up])


@pytest.mark.parametrize("n", [1, 2, 17, 64, 200])
@pytest.mark.parametrize("alpha,beta", [(1.0, 1.0), (2.5, 0.3), (0.8, -0.05), (0.1, 7.0)])
def test_update_matches_refactorisation(n, alpha, beta):
    """Closed form: the result is THE Cholesky factor of alpha A + beta v v^T
    (unique), so it equals LAPACK's and reproduces the updated matrix."""
    A, v = _update_case(n, alpha, beta)
    L = np.linalg.cholesky(A)
    U, info = orc.chol_update(L, alpha, beta, v, row_major=False)
    M = alpha * A + beta * np.outer(v, v)
    assert info == 0
    Lu = np.tril(U)
    assert np.max(np.abs(Lu - np.linalg.cholesky(M))) <= 1e-13 * np.max(np.abs(Lu))
    assert np.linalg.norm(Lu @ Lu.T - M) <= 1e-14 * n * np.linalg.norm(M)


def test_update_beta_zero_scales_whole_storage():
    """beta == 0 multiplies the whole stored matrix by sqrt(alpha)
    (decompositions.hpp:145-148)."""
    F = np.arange(16.0).reshape(4, 4) + 1
    U, info = orc.chol_update(F, 4.0, 0.0, np.ones(4))
    assert info == 0 and np.array_equal(U, 2.0 * F)


def test_update_indefinite_reports_first_failing_column():
    """A downdate that leaves alpha A + beta v v^T indefinite: info is the
    order of the first leading minor that is not positive definite (checked by
    brute force), and the columns before it hold that matrix's partial
    factor, as the reference leaves them when it throws."""
    n = 60
    A, v = _update_case(n, 1.0, -10.0, seed=3)
    M = A - 10.0 * np.outer(v, v)
    first = next(k for k in range(1, n + 1) if np.min(np.linalg.eigvalsh(M[:k, :k])) <= 0)
    U, info = orc.chol_update(np.linalg.cholesky(A), 1.0, -10.0, v, row_major=False)
    assert info == first
    j = info - 1
    Lp = np.linalg.cholesky(M[:j, :j]) if j else np.zeros((0, 0))
    # columns 0..j-1 of the factor of M: the leading block and the rows below
    below = np.linalg.solve(Lp, M[:j, j:]).T if j else np.zeros((n, 0))
    assert np.allclose(np.tril(U[:j, :j]), Lp, rtol=1e-11, atol=1e-12)
    assert np.allclose(U[j:, :j], below, rtol=1e-10, atol=1e-11)


@pytest.mark.parametrize("incv", [1, 3])
def test_update_strided_vector(incv):
    A, v = _update_case(30, 1.5, 0.4, seed=5)
    vs = np.full(30 * incv, np.nan)
    vs[::incv] = v
    L = np.linalg.cholesky(A)
    U1, _ = orc.chol_update(L, 1.5, 0.4, v, row_major=False)
    U2, _ = orc.chol_update(L, 1.5, 0.4, vs, row_major=False, incv=incv)
    assert np.array_equal(U1, U2)


# ---------------------------------------------------------------- pstrf --
def _perm_from_swaps(P):
    perm = np.arange(len(P))
    for i, p in enumerate(P):
        perm[[i, p]] = perm[[p, i]]
    return perm


@pytest.mark.parametrize("row_major", ORIENT)
@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("rank", [0, 50])
def test_pstrf_reference_fixture(rank, upper, row_major):
    """Test/cholesky.cpp:84-143: createSymm(123) has full rank and
    createSymm(123, 50) rank 50; L L^T (or U^T U) with the permutation undone
    reconstructs A to 1e-12."""
    n = 123
    A = fixtures.create_symm(n, rank)
    F, P, r = orc.pstrf(A, upper=upper, row_major=row_major)
    assert r == (rank or n)
    L = np.triu(F).T if upper else np.tril(F)
    assert np.max(np.abs(A - orc.swap_full_inverted(P, L @ L.T))) <= 1e-12
    # beyond the rank the factor is zero
    assert np.all(L[:, r:] == 0)


@pytest.mark.parametrize("n", [1, 7, 20, 21, 64, 150])
def test_pstrf_matches_lapack(n):
    """LAPACK dpstrf (scipy) picks the same pivots (largest remaining
    diagonal, first on ties) and the same factor on a matrix with well
    separated pivots."""
    from scipy.linalg import lapack
    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n))
    A = G @ G.T / n + np.diag(np.linspace(0.5, 3.0, n)[rng.permutation(n)])
    F, P, r = orc.pstrf(A, row_major=False)
    c, piv, rank, info = lapack.dpstrf(A, lower=1)
    assert r == rank == n
    assert np.array_equal(_perm_from_swaps(P), piv - 1)
    assert np.allclose(np.tril(F), np.tril(c), rtol=1e-12, atol=1e-12)


def test_pstrf_ties_take_the_first_maximum():
    """A = I: every pivot ties, std::max_element takes the first, so no row
    moves and L = I; diag(2, 3, 3, 1) pivots on the first 3."""
    F, P, r = orc.pstrf(np.eye(5))
    assert r == 5 and np.array_equal(P, np.arange(5)) and np.array_equal(np.tril(F), np.eye(5))
    F, P, r = orc.pstrf(np.diag([2.0, 3.0, 3.0, 1.0]))
    assert P[0] == 1 and r == 4


# ----------------------------------------------------------------- trsv --
TRSV_CASES = [(left, upper, unit) for left in (True, False) for upper in (False, True) for unit in (False, True)]


def _tri(A, upper, unit):
    T = np.triu(A) if upper else np.tril(A)
    if unit:
        T = T - np.diag(np.diag(T)) + np.eye(len(A))
    return T


@pytest.mark.parametrize("left,upper,unit", TRSV_CASES)
@pytest.mark.parametrize("a_row_major", ORIENT)
def test_trsv_reference_fixture(left, upper, unit, a_row_major):
    """Test/triangular_solve.cpp:19-110 (size 158): A x = b (left) or
    x A = b (right) to 1e-11, and LAPACK's trtrs (scipy) to rounding."""
    Alow = fixtures.triangular_matrix(158, False)
    A = Alow if not upper else Alow.T.copy()
    b = fixtures.triangular_vector_rhs(158)
    x, sing = orc.trsv(A, b, upper, unit, left, a_row_major=a_row_major)
    assert not sing
    T = _tri(A, upper, unit)
    r = T @ x - b if left else x @ T - b
    assert np.max(np.abs(r)) < 1e-11
    ref = sla.solve_triangular(T if left else T.T, b, lower=(not upper) if left else upper)
    np.testing.assert_allclose(x, ref, rtol=1e-13, atol=1e-15 * np.max(np.abs(ref)))


@pytest.mark.parametrize("left,upper,unit", TRSV_CASES)
@pytest.mark.parametrize("a_row_major", ORIENT)
def test_trsv_vs_mpmath(left, upper, unit, a_row_major):
    """40-digit solves of random triangular systems (n = 29): every variant
    of the reference's trsv_impl agrees to 1e-14 of the solution's size."""
    import mpmath
    mpmath.mp.dps = 40
    n = 29
    A = synth.uniform(3, 0, n * n, -1, 1).reshape(n, n) + 3 * np.eye(n)
    b = synth.uniform(3, 1, n, -1, 1)
    x, _ = orc.trsv(A, b, upper, unit, left, a_row_major=a_row_major)
    T = _tri(A, upper, unit)
    M = mpmath.matrix((T if left else T.T).tolist())
    xe = np.array([float(v) for v in mpmath.lu_solve(M, mpmath.matrix(b.tolist()))])
    assert np.max(np.abs(x - xe)) < 1e-14 * np.max(np.abs(xe))


@pytest.mark.parametrize("a_row_major", ORIENT)
@pytest.mark.parametrize("inc", [1, 3])
def test_trsv_vs_lapack_strided(a_row_major, inc):
    n = 400
    A = np.tril(synth.uniform(5, 0, n * n, -1, 1).reshape(n, n)) / np.sqrt(n) + 2 * np.eye(n)
    b = synth.uniform(5, 1, n, -1, 1)
    for upper, T in ((False, A), (True, A.T.copy())):
        x, _ = orc.trsv(T, b, upper, False, True, a_row_major=a_row_major, inc=inc)
        ref = sla.solve_triangular(T, b, lower=not upper)
        np.testing.assert_allclose(x, ref, rtol=0, atol=1e-13 * np.max(np.abs(ref)))


def test_trsv_exact_on_integers():
    """A unit lower triangle of small integers and an integer b: the solution
    is integral and every intermediate exact, so both storage orders (the
    column sweep and the row dot products) must give the exact answer."""
    from fractions import Fraction
    n = 14
    rng = np.random.default_rng(2)
    A = np.tril(rng.integers(-1, 2, (n, n)).astype(float), -1) + np.diag(rng.integers(2, 9, n).astype(float))
    b = rng.integers(-5, 6, n).astype(float)
    x_exact = []
    for i in range(n):  # unit solve in exact rationals
        x_exact.append(Fraction(int(b[i])) - sum(Fraction(int(A[i, k])) * x_exact[k] for k in range(i)))
    xe = np.array([float(v) for v in x_exact])
    for rm in ORIENT:
        x, _ = orc.trsv(A, b, False, True, True, a_row_major=rm)
        assert np.array_equal(x, xe)


def test_trsv_special_cases():
    n = 33
    b = synth.uniform(6, 1, n, -1, 1)
    for rm in ORIENT:
        x, _ = orc.trsv(np.eye(n), b, False, False, True, a_row_major=rm)
        assert np.array_equal(x, b)
        d = synth.uniform(6, 2, n, 1, 2)
        x, _ = orc.trsv(np.diag(d), b, True, False, True, a_row_major=rm)
        assert np.array_equal(x, b / d)
        # unit: the stored diagonal is never read
        G = np.tril(synth.uniform(6, 3, n * n, -1, 1).reshape(n, n)) / n
        Gd = G.copy()
        Gd[np.diag_indices(n)] = 7.0
        x1, _ = orc.trsv(G, b, False, True, True, a_row_major=rm)
        x2, _ = orc.trsv(Gd, b, False, True, True, a_row_major=rm)
        assert np.array_equal(x1, x2)
        # a zero diagonal: the reference throws "[TRSV] Matrix is singular!"
        Z = np.tril(np.ones((n, n)))
        Z[10, 10] = 0.0
        assert orc.trsv(Z, b, False, False, True, a_row_major=rm)[1]
        assert not orc.trsv(Z, b, False, True, True, a_row_major=rm)[1]


# ----------------------------------------------------------------- gemv --
@pytest.mark.parametrize("a_row_major", ORIENT)
def test_gemv_reference_fixture(a_row_major):
    """Test/hip_prod.cpp:33-106: result(1.5) += -2 A x for A 50 x 40 (and
    A^T x through the other orientation), against the textbook loop."""
    A, x = fixtures.prod_vector_operands()
    y = orc.gemv(A, x, np.full(50, 1.5), -2.0, a_row_major=a_row_major)
    ref = np.array([1.5 + sum(-2.0 * A[i, k] * x[k] for k in range(40)) for i in range(50)])
    np.testing.assert_allclose(y, ref, rtol=1e-14)
    xt = 0.1 * np.arange(50) + 0.1
    yt = orc.gemv(A.T.copy(), xt, np.zeros(40), 1.0, a_row_major=a_row_major)
    np.testing.assert_allclose(yt, A.T @ xt, rtol=1e-14)


def test_gemv_exact_on_integers():
    """Small integers, alpha = -2 and y = 1.5: every product and partial sum
    is exact, so both orientations (row dots, column axpys) equal numpy."""
    rng = np.random.default_rng(4)
    A = rng.integers(-9, 10, (70, 45)).astype(float)
    x = rng.integers(-9, 10, 45).astype(float)
    ref = 1.5 - 2.0 * (A @ x)
    for rm in ORIENT:
        assert np.array_equal(orc.gemv(A, x, np.full(70, 1.5), -2.0, a_row_major=rm), ref)


# ------------------------------------------------------ vector SPD solve --
@pytest.mark.parametrize("a_row_major", ORIENT)
def test_posv_vector_reference_fixture(a_row_major):
    """Test/symm_solve.cpp:89-139 restated for one right-hand side: A x = b
    to 1e-12; the vector path (trsv) agrees with the matrix path (trsm) on
    the same column."""
    A = fixtures.create_symm(128)
    b = fixtures.symm_solve_rhs(128, 151)[:, 17].copy()
    x, _, info, sing = orc.posv_vec(A, b, a_row_major=a_row_major)
    assert info == 0 and not sing
    assert np.max(np.abs(A @ x - b)) < 1e-12
    X, _, _, _ = orc.posv(A, b[:, None], a_row_major=a_row_major)
    np.testing.assert_allclose(x, X[:, 0], rtol=1e-13, atol=1e-15)


# -------------------------------------------------------------- mutants --
@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("mutants")


def _mutant(mutant_dir, name):
    path = os.path.join(str(mutant_dir), f"liboracle_{name}.so")
    if not os.path.exists(path):
        orc_build.build(path, mutant=name)
    return orc.load(path)


def test_mutant_ugemm_dropped_term_is_caught(mutant_dir):
    lib = _mutant(mutant_dir, "UGEMM_DROP_LAST_K")
    A, B = fixtures.prod_operands()
    got = orc.gemm(A, B, np.full((80, 50), 1.5), -2.0, lib=lib)
    assert not np.allclose(got, 1.5 - 2.0 * A @ B, rtol=1e-12)


def test_mutant_trsm_transposed_is_caught(mutant_dir):
    lib = _mutant(mutant_dir, "TRSM_TRANSPOSED_A")
    A = fixtures.triangular_matrix(158, True)
    B = fixtures.triangular_rhs(158, 138)
    X, _ = orc.trsm(A, B, False, False, True, lib=lib)
    assert np.max(np.abs(np.tril(A) @ X - B)) > 1e-11


def test_mutant_potrf_sign_is_caught(mutant_dir):
    lib = _mutant(mutant_dir, "POTRF_SIGN")
    A = fixtures.create_symm(123)
    F, info = orc.potrf(A, lib=lib)
    L = np.tril(F)
    assert info != 0 or not np.all(np.isfinite(L)) or np.max(np.abs(A - L @ L.T)) > 1e-12


def test_mutant_syrk_triangle_is_caught(mutant_dir):
    lib = _mutant(mutant_dir, "SYRK_WRONG_TRIANGLE")
    arg = fixtures.syrk_arg()
    M = orc.syrk(arg, np.full((384, 384), 3.0), 2.0, False, lib=lib)
    mask = np.tril(np.ones((384, 384), bool))
    assert not (np.allclose(M[mask], (2.0 * arg @ arg.T + 3.0)[mask], rtol=1e-12) and np.all(M[~mask] == 3.0))


def test_mutant_update_unscaled_temp_is_caught(mutant_dir):
    lib = _mutant(mutant_dir, "UPDATE_UNSCALED_TEMP")
    n = 123
    A = fixtures.create_symm(n)
    v = 0.1 / n * np.arange(n) - 0.5
    F, _ = orc.potrf(A)
    U, info = orc.chol_update(F, 0.1, 0.5, v, lib=lib)
    R, _ = orc.potrf(0.1 * A + 0.5 * np.outer(v, v))
    lo = np.tril_indices(n)
    assert info != 0 or not np.allclose(U[lo], R[lo], rtol=1e-13, atol=0)


def test_mutant_pstrf_last_max_is_caught(mutant_dir):
    lib = _mutant(mutant_dir, "PSTRF_LAST_MAX")
    F, P, r = orc.pstrf(np.eye(5), lib=lib)
    assert not np.array_equal(P, np.arange(5))


def test_mutant_trsv_skip_first_is_caught(mutant_dir):
    lib = _mutant(mutant_dir, "TRSV_SKIP_FIRST")
    A = fixtures.triangular_matrix(158, False)
    b = fixtures.triangular_vector_rhs(158) + 0.5  # the fixture's b(0) = 0 would hide a dropped first term
    x, _ = orc.trsv(A, b, False, False, True, a_row_major=True, lib=lib)
    assert np.max(np.abs(np.tril(A) @ x - b)) > 1e-11


def test_mutant_gemv_no_alpha_is_caught(mutant_dir):
    lib = _mutant(mutant_dir, "GEMV_NO_ALPHA")
    A, x = fixtures.prod_vector_operands()
    y = orc.gemv(A, x, np.full(50, 1.5), -2.0, a_row_major=False, lib=lib)
    assert not np.allclose(y, 1.5 - 2.0 * A @ x, rtol=1e-12)
