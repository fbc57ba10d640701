"""Cholesky adjoint on the GPU (rmh_?potrf_adjoint) through the C-ABI.

The kernel is wrapped in a torch.autograd.Function here, in the test harness
only.  Checks: the CPU FP64 reference adjoint (oracle/adjoint.py) on random
factors, both storage triangles; the Gaussian-process negative log marginal
likelihood 0.5 y^T A^-1 y + sum log L_ii, whose gradient has the closed form
0.5 (A^-1 - alpha alpha^T); and central finite differences (h = 1e-6) on
random symmetric directions.  Workload: RBF matrix of points in [0,1]^8,
lengthscale 0.3, + 1e-2 I, y ~ N(0, 1).
"""
import numpy as np
import pytest
import torch

from oracle import adjoint as adj
from remora_amd import capi, synth
from remora_amd import linalg as la

pytestmark = pytest.mark.gpu

_H = {}


def handle():
    if "h" not in _H:
        _H["h"] = capi.Handle(0)
    return _H["h"]


class RemoraCholesky(torch.autograd.Function):
    """L = chol(A) forward by rmh_dpotrf, backward by rmh_dpotrf_adjoint."""

    @staticmethod
    def forward(ctx, A):
        L = la.colmajor(A.detach().clone())
        info = la.potrf(handle(), L)
        assert int(info.item()) == 0
        L = torch.tril(L)
        ctx.save_for_backward(L)
        return L

    @staticmethod
    def backward(ctx, Lbar):
        (L,) = ctx.saved_tensors
        return la.potrf_adjoint(handle(), la.colmajor(L), la.colmajor(Lbar.contiguous()))


def nlml(A, y):
    L = RemoraCholesky.apply(A)
    z = torch.linalg.solve_triangular(L, y[:, None], upper=False)[:, 0]
    return 0.5 * torch.dot(z, z) + torch.log(torch.diagonal(L)).sum()


def gp_problem(n, seed=1):
    x = synth.points(n, 8, seed=seed)
    A = synth.rbf(x, 0.3, 1e-2)
    y = synth.normal(seed + 10, 0, n)
    return A, y


@pytest.mark.parametrize("n", [1, 17, 64, 130, 700])
@pytest.mark.parametrize("upper", [False, True])
def test_adjoint_vs_cpu_reference(n, upper):
    A = synth.gram_spd(n, seed=n)
    L = np.linalg.cholesky(A)
    rng = np.random.default_rng(n)
    Lbar = np.tril(rng.standard_normal((n, n)))
    ref = adj.potrf_adjoint(L, Lbar)
    # garbage in the unused triangle must be ignored
    Ls = L.T.copy() if upper else L.copy()
    Lbs = Lbar.T.copy() if upper else Lbar.copy()
    junk = np.triu(np.full((n, n), 7.0), 1)
    if upper:
        Ls += junk.T
        Lbs += junk.T
    else:
        Ls += junk
        Lbs += junk
    dL = la.colmajor(torch.from_numpy(Ls).cuda())
    dLb = la.colmajor(torch.from_numpy(Lbs).cuda())
    got = la.potrf_adjoint(handle(), dL, dLb, upper=upper).cpu().numpy()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10 * max(1.0, np.max(np.abs(ref))))
    assert np.array_equal(got, got.T)


def test_gp_nlml_gradient_closed_form():
    n = 512
    A, y = gp_problem(n)
    At = torch.tensor(A, device="cuda", requires_grad=True)
    f = nlml(At, torch.tensor(y, device="cuda"))
    f.backward()
    g = At.grad.cpu().numpy()
    ref = adj.gp_nlml_grad(A, y)
    assert abs(f.item() - adj.gp_nlml(A, y)) <= 1e-10 * abs(adj.gp_nlml(A, y))
    np.testing.assert_allclose(g, ref, rtol=0, atol=1e-8 * np.max(np.abs(ref)))


def test_gp_nlml_central_differences():
    n = 256
    A, y = gp_problem(n, seed=4)
    yt = torch.tensor(y, device="cuda")
    At = torch.tensor(A, device="cuda", requires_grad=True)
    nlml(At, yt).backward()
    g = At.grad.cpu().numpy()
    rng = np.random.default_rng(7)
    h = 1e-6
    for _ in range(32):
        V = rng.standard_normal((n, n))
        V = 0.5 * (V + V.T)
        with torch.no_grad():
            fp = nlml(torch.tensor(A + h * V, device="cuda"), yt).item()
            fm = nlml(torch.tensor(A - h * V, device="cuda"), yt).item()
        fd = (fp - fm) / (2 * h)
        an = float(np.sum(g * V))
        assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)), (fd, an)


def test_gp_nlml_gradient_n8192():
    """The differentiation workload at full size: closed form on the GPU."""
    n = 8192
    A, y = gp_problem(n, seed=1)
    h = handle()
    At = torch.tensor(A, device="cuda", requires_grad=True)
    yt = torch.tensor(y, device="cuda")
    nlml(At, yt).backward()
    g = At.grad
    # closed form with the library's own solve: A^-1 and alpha
    L = la.colmajor(torch.tensor(A, device="cuda"))
    la.potrf(h, L)
    Ainv = la.colmajor(torch.eye(n, dtype=torch.float64, device="cuda"))
    la.potrs(h, L, Ainv)
    alpha = Ainv @ yt
    ref = 0.5 * (Ainv - torch.outer(alpha, alpha))
    ref = 0.5 * (ref + ref.t())
    err = (g - ref).abs().max().item() / ref.abs().max().item()
    assert err <= 1e-8, err
    # four finite-difference directions at full size
    rng = np.random.default_rng(3)
    for _ in range(4):
        V = torch.tensor(rng.standard_normal((n, n)), device="cuda")
        V = 0.5 * (V + V.t())
        with torch.no_grad():
            fp = nlml(At.detach() + 1e-6 * V, yt).item()
            fm = nlml(At.detach() - 1e-6 * V, yt).item()
        fd = (fp - fm) / 2e-6
        an = float((g * V).sum().item())
        assert abs(fd - an) <= 1e-3 * max(1.0, abs(an)), (fd, an)


def test_adjoint_rejects_aliasing():
    L = la.colmajor(torch.eye(4, dtype=torch.float64, device="cuda"))
    with pytest.raises(capi.RemoraError):
        la.potrf_adjoint(handle(), L, L, Abar=L)
