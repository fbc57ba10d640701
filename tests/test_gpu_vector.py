"""GPU parity of the vector right-hand-side path against the oracle.

rmh_?trsv and rmh_?gemv (the layer under kernels::trsv / kernels::gemv on
hip_tag) and the vector SPD solve (cholesky_decomposition::solve(b): potrf,
then trsv<lower> and trsv<upper>) are compared with the oracle's restatement of
the reference's CPU kernels (default/trsv.hpp:44-183, default/gemv.hpp:43-93)
at the reference test's size (158) and at sizes where the device path takes
its blocked route (n >= 512: block inverses and, for one right-hand side, the
split GEMV of the trailing updates).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from remora_amd import synth
from remora_amd import linalg as la
from tests import fixtures
from tests.gpu_util import dev, host

pytestmark = pytest.mark.gpu

TRSV = [(upper, trans, unit) for upper in (False, True) for trans in (False, True) for unit in (False, True)]


def _vec(v, inc, dtype=torch.float64):
    """Device buffer holding v at stride inc (NaN between the elements)."""
    buf = torch.full((max(1, (len(v) - 1) * inc + 1),), float("nan"), dtype=dtype)
    buf[::inc] = torch.from_numpy(np.asarray(v, dtype=np.float64)).to(dtype)
    return buf.cuda()


def _get(buf, inc, n):
    torch.cuda.synchronize()
    out = buf.cpu().double().numpy()
    return out[::inc][:n].copy(), out


def _oracle_trsv(A, b, upper, trans, unit):
    # op(A) x = b with A column-major: trans is the right-side solve x A = b
    return orc.trsv(A, b, upper, unit, left=not trans, a_row_major=False)


def _tri_matrix(n, upper, seed):
    L = np.tril(synth.uniform(seed, 0, n * n, -1, 1).reshape(n, n)) / np.sqrt(n) + 2.0 * np.eye(n)
    return L.T.copy() if upper else L


@pytest.mark.parametrize("upper,trans,unit", TRSV)
@pytest.mark.parametrize("incx", [1, 3])
def test_dtrsv_reference_fixture_vs_oracle(handle, upper, trans, unit, incx):
    """Test/triangular_solve.cpp:19-110 (n = 158): the device solve equals the
    oracle's trsv_impl to 1e-13 of the solution; the gaps of a strided x are
    untouched."""
    Alow = fixtures.triangular_matrix(158, False)
    A = Alow if not upper else Alow.T.copy()
    b = fixtures.triangular_vector_rhs(158) + 0.25
    ref, sing = _oracle_trsv(A, b, upper, trans, unit)
    assert not sing
    dx = _vec(b, incx)
    la.trsv(handle, dev(A), dx, upper=upper, trans=trans, unit=unit, incx=incx)
    x, raw = _get(dx, incx, 158)
    np.testing.assert_allclose(x, ref, rtol=0, atol=1e-13 * np.max(np.abs(ref)))
    if incx > 1:
        mask = np.ones(len(raw), bool)
        mask[::incx] = False
        assert np.all(np.isnan(raw[mask]))


@pytest.mark.parametrize("upper,trans,unit", TRSV)
@pytest.mark.parametrize("n", [600, 4099, 8192])
def test_dtrsv_blocked_vs_oracle(handle, upper, trans, unit, n):
    """Orders at which the device runs the blocked solve (block inverses, split
    GEMV trailing updates): against the oracle to 1e-12 of the solution."""
    A = _tri_matrix(n, upper, seed=n)
    b = synth.uniform(n, 1, n, -1, 1)
    ref, _ = _oracle_trsv(A, b, upper, trans, unit)
    dx = _vec(b, 1)
    la.trsv(handle, dev(A, ld=n + 3), dx, upper=upper, trans=trans, unit=unit)
    x, _ = _get(dx, 1, n)
    np.testing.assert_allclose(x, ref, rtol=0, atol=1e-12 * np.max(np.abs(ref)))


@pytest.mark.parametrize("upper,trans", [(False, False), (True, True)])
def test_strsv_vs_oracle(handle, upper, trans):
    """FP32 at n = 2000 against the FP64 oracle, to FP32 rounding."""
    n = 2000
    A = _tri_matrix(n, upper, seed=7)
    b = synth.uniform(7, 1, n, -1, 1)
    A32 = A.astype(np.float32).astype(np.float64)
    b32 = b.astype(np.float32).astype(np.float64)
    ref, _ = _oracle_trsv(A32, b32, upper, trans, False)
    dx = _vec(b, 1, torch.float32)
    la.trsv(handle, dev(A, torch.float32), dx, upper=upper, trans=trans)
    x, _ = _get(dx, 1, n)
    np.testing.assert_allclose(x, ref, rtol=0, atol=2e-5 * np.max(np.abs(ref)))


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("m,n,incx,incy", [(50, 40, 1, 1), (700, 333, 2, 3), (8192, 8192, 1, 1), (3001, 5003, 1, 2)])
def test_dgemv_vs_oracle(handle, trans, m, n, incx, incy):
    """y = alpha op(A) x + y with alpha = -2 and y = 1.5 (Test/hip_prod.cpp's
    cases) against the oracle's gemv_impl, componentwise to 1e-13 of
    |alpha| |op(A)| |x| + |y|; strided x and y."""
    if (m, n) == (50, 40):
        A, xv = fixtures.prod_vector_operands()
        if trans:
            xv = 0.1 * np.arange(50) + 0.1
    else:
        A = synth.normal(m + n, 0, m * n).reshape(m, n)
        xv = synth.normal(m + n, 1, m if trans else n)
    ylen = n if trans else m
    y0 = np.full(ylen, 1.5)
    opA = A.T.copy() if trans else A
    ref = orc.gemv(opA, xv, y0, -2.0, a_row_major=False)
    dy = _vec(y0, incy)
    la.gemv(handle, dev(A), _vec(xv, incx), dy, alpha=-2.0, beta=1.0, trans=trans, incx=incx, incy=incy)
    y, raw = _get(dy, incy, ylen)
    scale = 2.0 * np.abs(opA) @ np.abs(xv) + 1.5
    assert np.all(np.abs(y - ref) <= 1e-13 * scale)
    if incy > 1:
        mask = np.ones(len(raw), bool)
        mask[::incy] = False
        assert np.all(np.isnan(raw[mask]))


def test_dgemv_beta_scales_y(handle):
    """beta != 1 (blas_device::gemv's beta): y = alpha A x + beta y equals the
    oracle's y += alpha A x applied to beta y."""
    m, n = 900, 700
    A = synth.normal(3, 0, m * n).reshape(m, n)
    xv = synth.normal(3, 1, n)
    y0 = synth.normal(3, 2, m)
    ref = orc.gemv(A, xv, 0.25 * y0, 0.5, a_row_major=False)
    dy = _vec(y0, 1)
    la.gemv(handle, dev(A), _vec(xv, 1), dy, alpha=0.5, beta=0.25)
    y, _ = _get(dy, 1, m)
    scale = 0.5 * np.abs(A) @ np.abs(xv) + 0.25 * np.abs(y0)
    assert np.all(np.abs(y - ref) <= 1e-13 * scale)


def test_sgemv_reference_fixture(handle):
    """Test/hip_prod.cpp:57-106 in float: 1e-2 percent of the textbook value."""
    A, xv = fixtures.prod_vector_operands()
    ref = orc.gemv(A, xv, np.full(50, 1.5), -2.0)
    dy = _vec(np.full(50, 1.5), 1, torch.float32)
    la.gemv(handle, dev(A, torch.float32), _vec(xv, 1, torch.float32), dy, alpha=-2.0, beta=1.0)
    y, _ = _get(dy, 1, 50)
    np.testing.assert_allclose(y, ref, rtol=1e-4)


@pytest.mark.parametrize("n,upper", [(128, False), (128, True), (5000, False), (5000, True)])
def test_vector_spd_solve_vs_oracle(handle, n, upper):
    """cholesky_decomposition::solve(b) (decompositions.hpp:111-115) through
    the ABI -- potrf, trsv(L), trsv(L^T) -- against the oracle's restatement of
    the same sequence: createSymm(128) (Test/symm_solve.cpp:89-139) and a
    Gram matrix at n = 5000."""
    if n == 128:
        A = fixtures.create_symm(128)
        b = fixtures.symm_solve_rhs(128, 151)[:, 17].copy()
    else:
        A = synth.gram_spd(n, seed=17)
        b = synth.uniform(17, 1, n, -1, 1)
    xo, _, info, sing = orc.posv_vec(A, b, a_row_major=False)
    assert info == 0 and not sing
    dA = dev(A)
    assert int(la.potrf(handle, dA, upper=upper).item()) == 0
    dx = _vec(b, 1)
    # lower storage: L y = b, L^T x = y; upper storage (U = L^T): U^T y = b, U x = y
    la.trsv(handle, dA, dx, upper=upper, trans=upper)
    la.trsv(handle, dA, dx, upper=upper, trans=not upper)
    x, _ = _get(dx, 1, n)
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-12 * np.max(np.abs(xo)))
    assert synth.backward_error(A, x[:, None], b[:, None]) < 1e-15


@pytest.mark.parametrize("upper,trans,unit", TRSV)
@pytest.mark.parametrize("n,incx", [(37, 1), (1000, 2), (4099, 1)])
def test_dtrmv_exact_on_integers(handle, upper, trans, unit, n, incx):
    """x = op(T) x (blas_device::trmv).  Small integers make every product and
    partial sum exact in FP64, so the device result must equal op(T) x
    exactly; the stored other triangle (set to NaN) and, with unit, the stored
    diagonal (set to 7) are never read."""
    rng = np.random.default_rng(n + 8 * upper + 4 * trans + 2 * unit)
    A = rng.integers(-4, 5, (n, n)).astype(float)
    T = np.triu(A) if upper else np.tril(A)
    S = T.copy()
    S[np.tril_indices(n, -1) if upper else np.triu_indices(n, 1)] = np.nan
    if unit:
        T[np.diag_indices(n)] = 1.0
        S[np.diag_indices(n)] = 7.0
    xv = rng.integers(-4, 5, n).astype(float)
    ref = (T.T if trans else T) @ xv
    dx = _vec(xv, incx)
    la.trmv(handle, dev(S), dx, upper=upper, trans=trans, unit=unit, incx=incx)
    x, _ = _get(dx, incx, n)
    assert np.array_equal(x, ref)


def test_strmv_rounding(handle):
    n = 3000
    A = _tri_matrix(n, False, seed=9)
    xv = synth.uniform(9, 1, n, -1, 1)
    ref = np.tril(A) @ xv
    dx = _vec(xv, 1, torch.float32)
    la.trmv(handle, dev(A, torch.float32), dx)
    x, _ = _get(dx, 1, n)
    assert np.max(np.abs(x - ref)) <= 1e-5 * np.max(np.abs(np.tril(A)) @ np.abs(xv))
