"""Mixed-precision solve (rmh_dsposv): FP32-MFMA factor + FP64 refinement.

Pins: the backward error definition itself (<= 1e-14 on exit, recomputed
here on the host in FP64), agreement with the FP64 oracle solve to
~cond * 1e-15, and the LAPACK dsposv fallback contract."""
import numpy as np
import pytest

from oracle import oracle as orc
from remora_amd import synth
from tests.gpu_util import dev, host

pytestmark = pytest.mark.gpu


def _rbf_problem(n, nrhs, d=8, ls=0.3, jitter=1e-2, seed=1):
    X = synth.points(n, d, seed)
    A = synth.rbf(X, ls, jitter)
    B = synth.rhs(n, nrhs, seed=seed)
    return A, B


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("n,nrhs", [(1000, 8), (2600, 64)])
def test_dsposv_fp32_converges(handle, upper, n, nrhs):
    from remora_amd import linalg as la
    A, B = _rbf_problem(n, nrhs)
    dA = dev(A if not upper else A, ld=n + 2)
    # only the selected triangle is read: poison the other one
    import torch
    poison = torch.triu(torch.full((n, n), float("nan"), dtype=torch.float64, device="cuda"), 1)
    if upper:
        poison = poison.t()
    dA[:, :].masked_fill_(torch.isnan(poison), float("nan"))
    dB = dev(B)
    dX = dev(np.zeros((n, nrhs)))
    it, berr = la.dsposv(handle, dA, dB, dX, upper=upper, tol=1e-14, max_iter=30)
    X = host(dX)
    assert it >= 1, it
    assert berr <= 1e-14
    assert synth.backward_error(A, X, B) <= 2e-14
    Xref = np.linalg.solve(A, B)
    cond = np.linalg.cond(A)
    assert np.max(np.abs(X - Xref)) <= cond * 1e-15 * np.max(np.abs(Xref))


def test_dsposv_matches_fp64_oracle_small(handle):
    from remora_amd import linalg as la
    A, B = _rbf_problem(300, 4, ls=0.5, jitter=1e-3)
    dX = dev(np.zeros((300, 4)))
    it, berr = la.dsposv(handle, dev(A), dev(B), dX)
    Xo, _, info, _ = orc.posv(A, B)
    assert info == 0 and berr <= 1e-14
    np.testing.assert_allclose(host(dX), Xo, rtol=1e-8, atol=1e-10 * np.max(np.abs(Xo)))


def test_dsposv_fallback_when_fp32_factor_fails(handle):
    """PD in FP64 but not after rounding to FP32: dsposv returns the FP64
    solution with iters = -3."""
    from remora_amd import linalg as la
    n = 400
    G = synth.normal(4, 0, n * 40).reshape(n, 40)
    A = G @ G.T + 1e-9 * np.eye(n)   # rank 40 + tiny shift: cond ~1e12
    B = synth.rhs(n, 2, seed=4)
    dX = dev(np.zeros((n, 2)))
    it, berr = la.dsposv(handle, dev(A), dev(B), dX, max_iter=5)
    assert it < 0
    assert berr < 1e-14


def test_dsposv_not_positive_definite(handle):
    from remora_amd import capi, linalg as la
    A = -np.eye(50)
    with pytest.raises(capi.RemoraError) as e:
        la.dsposv(handle, dev(A), dev(np.ones((50, 1))), dev(np.zeros((50, 1))))
    assert e.value.status == capi.NOT_POSITIVE_DEFINITE


def test_spotrf_bf16x3_trailing_accuracy(handle):
    """The split-BF16x3 trailing update keeps about 16 bits per product:
    the FP32 factor error grows from ~1e-7 to ~1e-5, no more.  n = 8192
    reaches the 128 x 128 trailing tiles that use it."""
    from remora_amd import linalg as la
    import torch
    n = 8192
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    G = torch.randn((n, n), generator=g, dtype=torch.float64, device="cuda")
    A = (G @ G.t()) / n
    del G
    A.diagonal().add_(0.1)
    errs = {}
    for prec in ("fp32", "bf16x3"):
        la.set_spotrf_precision(handle, prec)
        L = la.colmajor(A.to(torch.float32))
        info = la.potrf(handle, L)
        assert int(info.item()) == 0
        Lt = torch.tril(L).double()
        errs[prec] = float((torch.linalg.norm(A - Lt @ Lt.t()) / torch.linalg.norm(A)).item())
        del L, Lt
    la.set_spotrf_precision(handle, "fp32")
    assert errs["fp32"] < 1e-6, errs
    assert 2 * errs["fp32"] < errs["bf16x3"] < 1e-4, errs


@pytest.mark.parametrize("n", [1000, 4096])
def test_dsposv_bf16x3(handle, n):
    """BF16x3 factor + FP64 refinement reaches 1e-14 (or falls back to FP64,
    LAPACK-style) on the RBF workload."""
    from remora_amd import linalg as la
    A, B = _rbf_problem(n, 16)
    dX = dev(np.zeros((n, 16)))
    it, berr = la.dsposv(handle, dev(A), dev(B), dX, prec="bf16x3", tol=1e-14, max_iter=30)
    assert berr <= 1e-14
    assert synth.backward_error(A, host(dX), B) <= 2e-14
    if it >= 0:
        # more corrections than the FP32 factor needs on the same problem
        dX2 = dev(np.zeros((n, 16)))
        it32, _ = la.dsposv(handle, dev(A), dev(B), dX2, prec="fp32")
        assert it >= it32


def test_dsposv_fp32_large_blocks(handle):
    """n = 9000 (> 8192): the FP32 refinement solves run on 2048-row blocks
    with a ragged last block; the result still reaches 1e-14."""
    from remora_amd import linalg as la
    n, nrhs = 9000, 64
    A, B = _rbf_problem(n, nrhs)
    dX = dev(np.zeros((n, nrhs)))
    it, berr = la.dsposv(handle, dev(A), dev(B), dX, tol=1e-14, max_iter=30)
    assert it >= 1, it
    assert berr <= 1e-14
    assert synth.backward_error(A, host(dX), B) <= 2e-14
