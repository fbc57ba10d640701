"""GPU parity of rmh_?potrf / rmh_?potrs against the CPU oracle."""
import numpy as np
import pytest

from oracle import oracle as orc
from remora_amd import synth
from tests import fixtures
from tests.gpu_util import dev, host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("row_major", [True, False])
def test_potrf_reference_hip_fixture(handle, upper, row_major):
    """Test/hip_cholesky.cpp:36-65: createSymm(123); the GPU factor equals
    the CPU kernels::potrf element-wise -- untouched triangle included -- at
    1e-12 percent.  A row-major Remora matrix is the column-major transpose,
    so potrf<lower> on it is the ABI's upper factorisation."""
    from remora_amd import linalg as la
    A = fixtures.create_symm(123, hip_variant=True)
    ref, info = orc.potrf(A, upper=upper, row_major=row_major)
    assert info == 0
    abi_upper = upper != row_major  # row-major storage flips the triangle
    stored = A if not row_major else A.T.copy()     # column-major buffer content
    dA = dev(stored)
    dinfo = la.potrf(handle, dA, upper=abi_upper)
    got = host(dA)
    if row_major:
        got = got.T
    assert int(dinfo.item()) == 0
    np.testing.assert_allclose(got, ref, rtol=1e-14, atol=0)


@pytest.mark.parametrize("upper", [False, True])
def test_potrf_reference_cholesky_fixture(handle, upper):
    """Test/cholesky.cpp:40-80: the other triangle is untouched, A = L L^T to
    1e-12 (max abs), no NaN."""
    from remora_amd import linalg as la
    A = fixtures.create_symm(123)
    dA = dev(A)
    la.potrf(handle, dA, upper=upper)
    got = host(dA)
    off = np.triu(np.ones((123, 123), bool), 1)
    if upper:
        off = off.T
    assert np.array_equal(got[off], A[off])
    F = np.triu(got) if upper else np.tril(got)
    R = F.T @ F if upper else F @ F.T
    assert np.max(np.abs(A - R)) < 1e-12
    assert np.all(np.isfinite(got))


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 255, 257, 600, 1500])
@pytest.mark.parametrize("upper", [False, True])
def test_potrf_gram_vs_oracle(handle, n, upper):
    """A = G G^T/n + 0.1 I: GPU factor within 1e-12 relative of the FP64
    oracle (componentwise against |L||L^T|-scaled entries) and
    ||A - L L^T||_F / ||A||_F at the level of n*u."""
    from remora_amd import linalg as la
    A = synth.gram_spd(n, seed=n)
    ref, info = orc.potrf(A, upper=False, row_major=False)
    assert info == 0
    Lref = np.tril(ref)
    dA = dev(A, ld=n + 1)
    dinfo = la.potrf(handle, dA, upper=upper)
    got = host(dA)
    assert int(dinfo.item()) == 0
    L = np.triu(got).T if upper else np.tril(got)
    scale = np.sqrt(np.outer(np.diag(A), np.diag(A))) * np.tril(np.ones((n, n)))
    assert np.max(np.abs(L - Lref)) <= 1e-12 * max(1.0, np.max(scale))
    assert synth.factor_error(A, L) < 4 * n * np.finfo(float).eps


@pytest.mark.parametrize("nb,nbi", [(64, 64), (128, 32), (256, 64), (512, 64)])
def test_potrf_blocking_variants(handle, nb, nbi):
    from remora_amd import linalg as la
    n = 1100
    A = synth.gram_spd(n, seed=3)
    la.set_potrf_blocking(handle, nb, nbi)
    try:
        dA = dev(A)
        la.potrf(handle, dA)
        L = np.tril(host(dA))
    finally:
        la.set_potrf_blocking(handle, 0, 0)
    assert synth.factor_error(A, L) < 4 * n * np.finfo(float).eps


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("leaf", [32, 64, 128, 384, 4096])
def test_potrf_recursive_panel_leaves(handle, leaf, fp32):
    """Panels are halved recursively down to leaves of `leaf` columns (a
    leaf as wide as the panel is a single chain): every leaf width gives an
    n*u-accurate factor, and the look-ahead schedule is bitwise identical to
    the single stream for each."""
    import torch
    from remora_amd import linalg as la
    n = 3100
    A = synth.gram_spd(n, seed=17)
    la.set_potrf_blocking(handle, 1024, 32)
    la.set_potrf_leaf(handle, leaf)
    outs = []
    try:
        for mode in (1, 0):
            la.set_potrf_lookahead(handle, mode)
            dA = dev(A.astype(np.float32) if fp32 else A, torch.float32 if fp32 else torch.float64)
            info = la.potrf(handle, dA)
            assert int(info.item()) == 0
            outs.append(np.tril(host(dA)).astype(np.float64))
    finally:
        la.set_potrf_lookahead(handle, -1)
        la.set_potrf_leaf(handle, 0)
        la.set_potrf_blocking(handle, 0, 0)
    assert np.array_equal(outs[0], outs[1])
    tol = (4 * n * np.finfo(np.float32).eps) if fp32 else (4 * n * np.finfo(float).eps)
    assert synth.factor_error(A, outs[0]) < tol


@pytest.mark.parametrize("k", [0, 5, 64, 100, 700])
def test_potrf_info_not_positive_definite(handle, k):
    """LAPACK info: k+1 when the leading minor of order k+1 is the first that
    is not positive definite (the CPU reference returns nonzero there too)."""
    from remora_amd import linalg as la
    n = 800
    A = synth.gram_spd(n, seed=1)
    # make the leading (k+1) minor indefinite: scale down its last diagonal
    # entry below the Schur complement of the block before it
    L = np.linalg.cholesky(A[:k + 1, :k + 1])
    A[k, k] -= L[k, k] ** 2 * 1.5
    dA = dev(A)
    dinfo = la.potrf(handle, dA)
    assert int(dinfo.item()) == k + 1
    _, cpu_info = orc.potrf(A, row_major=False)
    assert cpu_info != 0


def test_potrf_nan_reports_info(handle):
    from remora_amd import linalg as la
    A = synth.gram_spd(200, seed=4)
    A[70, 70] = np.nan
    dinfo = la.potrf(handle, dev(A))
    assert int(dinfo.item()) == 71


def test_spotrf(handle):
    import torch
    from remora_amd import linalg as la
    n = 1000
    A = synth.gram_spd(n, seed=7)
    dA = dev(A, torch.float32)
    info = la.potrf(handle, dA)
    assert int(info.item()) == 0
    L = np.tril(host(dA))
    assert synth.factor_error(A, L) < 1e-5


@pytest.mark.parametrize("left", [True, False])
def test_potrs_reference_symm_solve(handle, left):
    """Test/symm_solve.cpp:89-139 (n=128, k=151): |A X - B| < 1e-12."""
    from remora_amd import linalg as la
    A = fixtures.create_symm(128)
    B = fixtures.symm_solve_rhs(128, 151)
    dA = dev(A)
    la.potrf(handle, dA)
    if left:
        dB = dev(B)
        la.potrs(handle, dA, dB)
        X = host(dB)
        err = np.max(np.abs(A @ X - B))
        ref, _, info, _ = orc.posv(A, B, left=True)
    else:
        # X A = B^T  <=>  A X^T = B  (A symmetric)
        Bt = B.T.copy()
        dB = dev(Bt.T.copy())
        la.potrs(handle, dA, dB)
        X = host(dB).T
        err = np.max(np.abs(X @ A - Bt))
        ref, _, info, _ = orc.posv(A, Bt, left=False)
    assert info == 0
    assert err < 1e-12
    np.testing.assert_allclose(X, ref, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("n,nrhs", [(512, 16), (2049, 64), (3000, 1)])
def test_potrf_potrs_backward_error(handle, upper, n, nrhs):
    from remora_amd import linalg as la
    A = synth.gram_spd(n, seed=0, divisor=512 if n == 512 else None)
    B = synth.rhs(n, nrhs, seed=0)
    dA = dev(A)
    la.potrf(handle, dA, upper=upper)
    dB = dev(B)
    la.potrs(handle, dA, dB, upper=upper)
    X = host(dB)
    assert synth.backward_error(A, X, B) < 1e-15
    if n <= 2049:
        Xref, _, _, _ = orc.posv(A, B)
        np.testing.assert_allclose(X, Xref, rtol=1e-10, atol=1e-12 * np.max(np.abs(Xref)))


@pytest.mark.parametrize("upper", [False, True])
def test_potrf_lookahead_is_bitwise_identical(handle, upper):
    """Look-ahead only reorders independent launches across two streams: the
    factor must be bit-identical to the single-stream schedule."""
    from remora_amd import linalg as la
    n = 4500  # large enough for 128 x 128 trailing tiles (mode 2 pads their LDS)
    A = synth.gram_spd(n, seed=11)
    outs = []
    for on in (1, 2, 0):
        la.set_potrf_lookahead(handle, on)
        try:
            dA = dev(A)
            info = la.potrf(handle, dA, upper=upper)
            outs.append(host(dA))
            assert int(info.item()) == 0
        finally:
            la.set_potrf_lookahead(handle, -1)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("upper", [False, True])
def test_potrf_cu_reserve_is_bitwise_identical(handle, upper, fp32):
    """Reserving compute units moves panel steps' trailing updates to a
    CU-masked stream: every step, a window of them, or none.  The launches
    and their order on each column range are unchanged, so the factor is
    bit-identical in every case.  With n = 5200 and panels of 1024 the steps
    have n - k1 = 4176, 3152, 2128, 1104 trailing rows, so the window
    (2500, 3500] moves one step to the masked stream and back."""
    from remora_amd import linalg as la
    n = 5200
    A = synth.gram_spd(n, seed=13)
    outs = []
    try:
        for cus, lo, hi in ((0, 0, 0), (4, 0, 1 << 62), (2, 2500, 3500), (4, 4096, 12288)):
            la.set_potrf_reserve(handle, cus, lo, hi)
            dA = dev(A.astype(np.float32) if fp32 else A)
            info = la.potrf(handle, dA, upper=upper)
            assert int(info.item()) == 0
            outs.append(host(dA))
    finally:
        la.set_potrf_reserve(handle, 0)
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)
    tol = (4 * n * np.finfo(np.float32).eps) if fp32 else (4 * n * np.finfo(float).eps)
    L = np.triu(outs[0]).T if upper else np.tril(outs[0])
    assert synth.factor_error(A, L.astype(np.float64)) < tol
    with pytest.raises(Exception):
        la.set_potrf_reserve(handle, 5)
    with pytest.raises(Exception):
        la.set_potrf_reserve(handle, -1)
    with pytest.raises(Exception):
        la.set_potrf_reserve(handle, 2, 100, 10)


@pytest.mark.parametrize("fp32", [False, True])
def test_potrf_narrow_panels(handle, fp32):
    """Modes 3 and 4 run the panel on four wavefronts (one K group): the two
    schedules give the same factor bit for bit, and it is as accurate as the
    eight-wavefront factor."""
    from remora_amd import linalg as la
    n = 4500
    A = synth.gram_spd(n, seed=12)
    outs = []
    for mode in (3, 4, 1):
        la.set_potrf_lookahead(handle, mode)
        try:
            dA = dev(A.astype(np.float32) if fp32 else A)
            info = la.potrf(handle, dA)
            outs.append(np.tril(host(dA)).astype(np.float64))
            assert int(info.item()) == 0
        finally:
            la.set_potrf_lookahead(handle, -1)
    assert np.array_equal(outs[0], outs[1])
    tol = (4 * n * np.finfo(np.float32).eps) if fp32 else (4 * n * np.finfo(float).eps)
    e_narrow, e_wide = synth.factor_error(A, outs[0]), synth.factor_error(A, outs[2])
    assert e_narrow < tol and e_narrow < 2 * e_wide + 1e-300, (e_narrow, e_wide)


@pytest.mark.parametrize("nbi", [16, 32, 64])
def test_potrf_large_lookahead_factor_error(handle, nbi):
    """Large n with look-ahead: panel kernels run beside the trailing update,
    so late workgroups of a panel launch are common -- the factor must stay
    exact to n*u (guards the diagonal-block write-back ordering)."""
    from remora_amd import linalg as la
    n = 6000
    A = synth.gram_spd(n, seed=21)
    la.set_potrf_blocking(handle, 256, nbi)
    try:
        dA = dev(A)
        info = la.potrf(handle, dA)
        L = np.tril(host(dA))
    finally:
        la.set_potrf_blocking(handle, 0, 0)
    assert int(info.item()) == 0
    assert synth.factor_error(A, L) < 4 * n * np.finfo(float).eps
