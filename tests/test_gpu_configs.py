"""Every workload size the benchmark configuration names, on the GPU.

The smaller tests compare the library with the oracle entry by entry at sizes
the oracle finishes in seconds.  Here the same entry points run at their full
sizes, each checked by what does not depend on the size or on this library:

* plain FP64 potrf + potrs at n = 4096 against the oracle (factor and solution
  componentwise), and at n = 16384 and 32768 by the FP64 backward error and
  the factor residual on sampled columns, both computed with torch (rocBLAS),
  not with this library's own routines, plus the untouched triangle, bit for
  bit (Test/hip_cholesky.cpp:36-65 and Test/cholesky.cpp:40-80 are the
  pattern);
* GEMM 8192^3 and SYRK n = k = 8192: a 64-row slab against the oracle's
  kernels::gemm restatement, a Freivalds probe A(Bx) against (AB)x over the
  whole product, and SYRK's other triangle bit for bit;
* mixed precision at n = 16384 (FP32 and split-BF16x3 trailing updates): the
  backward error <= 1e-14, recomputed here, and the solution against the
  FP64 path to cond(A) * 1e-15, with cond(A) bounded by ||A||_inf / 1e-2 (the
  RBF matrix is positive semi-definite plus 1e-2 I);
* the batched configuration: 4096 systems, n_i uniform in [16, 512], FP64 and
  FP32, every system's backward error;
* the adjoint at n = 8192: 32 central-difference directions (h = 1e-6) and the
  closed form 0.5 (A^-1 - alpha alpha^T) from torch.linalg.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from remora_amd import synth
from remora_amd import linalg as la
from tests.gpu_util import dev, host

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def _say(msg):
    print(msg, flush=True)


def gram_gpu(n, seed=0):
    """A = G G^T / n + 0.1 I, G N(0,1) from a seeded torch generator, built on
    the GPU with torch's own GEMM; column-major."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    G = torch.randn((n, n), generator=g, dtype=torch.float64, device="cuda")
    A = G @ G.t()
    del G
    A.div_(n)
    A = 0.5 * (A + A.t())
    A.diagonal().add_(0.1)
    return la.colmajor(A)


def rbf_gpu(n, d=8, lengthscale=0.3, jitter=1e-2, seed=1):
    x = torch.from_numpy(synth.points(n, d, seed)).cuda()
    sq = (x * x).sum(dim=1)
    d2 = (sq[:, None] + sq[None, :] - 2.0 * (x @ x.t())).clamp_min_(0.0)
    A = torch.exp(d2.mul_(-1.0 / (2.0 * lengthscale * lengthscale)))
    del d2
    A = 0.5 * (A + A.t())
    A.diagonal().add_(jitter)
    return la.colmajor(A)


def eta_torch(A, X, B):
    """max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf) in FP64
    with torch's GEMM (independent of this library)."""
    R = B - A @ X
    an = A.abs().sum(dim=1).max()
    return float((R.abs().amax(dim=0) / (an * X.abs().amax(dim=0) + B.abs().amax(dim=0))).max().item())


def factor_residual_sampled(A, L, cols=96, seed=5):
    """||A[:, S] - L L[S, :]^T||_F / ||A[:, S]||_F over sampled columns S."""
    n = A.shape[0]
    g = torch.Generator(device="cpu").manual_seed(seed)
    idx = torch.randperm(n, generator=g)[:cols].cuda()
    Lt = torch.tril(L)
    Ac = A[:, idx]
    return float(((Ac - Lt @ Lt[idx, :].t()).norm() / Ac.norm()).item())


# ------------------------------------------------------------- plain FP64 --
@pytest.mark.parametrize("upper", [False, True])
def test_p64_n4096_vs_oracle(handle, upper):
    """P64 at n = 4096 with 64 RHS: factor and solution against the oracle's
    kernels::potrf / cholesky_decomposition::solve restatement, componentwise."""
    n, nrhs = 4096, 64
    _say(f"p64 n={n} upper={upper}")
    A = synth.gram_spd(n, seed=0)
    B = synth.rhs(n, nrhs, seed=0)
    Xo, Lo, info, _ = orc.posv(A, B, a_row_major=False, b_row_major=False)
    assert info == 0
    Lref = np.tril(Lo)
    dA = dev(A)
    dinfo = la.potrf(handle, dA, upper=upper)
    dB = dev(B)
    la.potrs(handle, dA, dB, upper=upper)
    F = host(dA)
    X = host(dB)
    assert int(dinfo.item()) == 0
    L = np.triu(F).T if upper else np.tril(F)
    # |L - Lref|_ij <= 1e-12 * sqrt(A_ii A_jj) (|L| |L^T| bounds each entry of L by that)
    scale = np.sqrt(np.outer(np.diag(A), np.diag(A)))
    assert np.all(np.abs(L - Lref) <= 1e-12 * scale)
    # the other triangle is untouched
    off = np.tril(np.ones((n, n), bool), -1) if upper else np.triu(np.ones((n, n), bool), 1)
    assert np.array_equal(F[off], A[off])
    # solution: |X - Xo| <= 1e-12 * max_i |Xo(:, j)| per column
    assert np.all(np.abs(X - Xo) <= 1e-12 * np.max(np.abs(Xo), axis=0))
    assert synth.backward_error(A, X, B) < 1e-15
    assert synth.factor_error(A, L) < 4 * n * EPS


@pytest.mark.parametrize("n,upper", [(16384, False), (16384, True), (32768, False)])
def test_p64_full_size_properties(handle, n, upper):
    """P64 at n = 16384 and 32768 (8.6 GB): info 0, the FP64 backward error
    and sampled factor residual at the level of u, the other triangle bit for
    bit unchanged, a positive diagonal."""
    nrhs = 64
    _say(f"p64 n={n} upper={upper}")
    A = gram_gpu(n, seed=n)
    g = torch.Generator(device="cuda").manual_seed(n + 1)
    B = la.colmajor(torch.rand((n, nrhs), generator=g, dtype=torch.float64, device="cuda") * 2 - 1)
    F = la.colmajor(A.clone())
    info = la.potrf(handle, F, upper=upper)
    X = la.colmajor(B.clone())
    la.potrs(handle, F, X, upper=upper)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    if upper:
        assert torch.equal(torch.tril(F, -1), torch.tril(A, -1))
        L = torch.triu(F).t()
    else:
        assert torch.equal(torch.triu(F, 1), torch.triu(A, 1))
        L = torch.tril(F)
    del F
    assert bool((torch.diagonal(L) > 0).all())
    eta = eta_torch(A, X, B)
    res = factor_residual_sampled(A, L)
    _say(f"  eta {eta:.2e}  sampled ||A - L L^T||/||A|| {res:.2e}")
    assert eta < 1e-15
    assert res < 8 * EPS * np.sqrt(n)


# ------------------------------------------------------- GEMM / SYRK 8192 --
def test_dgemm_8192_cubed(handle):
    """C = A B + C (the reference's noalias(C) += prod(A, B)) at 8192^3 with
    A, B N(0,1): rows 4000..4063 against the oracle's kernels::gemm, and the
    whole product by a Freivalds probe."""
    n = 8192
    _say("dgemm 8192^3")
    g = torch.Generator(device="cuda").manual_seed(81)
    A = la.colmajor(torch.randn((n, n), generator=g, dtype=torch.float64, device="cuda"))
    Bm = la.colmajor(torch.randn((n, n), generator=g, dtype=torch.float64, device="cuda"))
    C0 = la.colmajor(torch.randn((n, n), generator=g, dtype=torch.float64, device="cuda"))
    C = la.colmajor(C0.clone())
    la.gemm(handle, A, Bm, C, alpha=1.0, beta=1.0)
    torch.cuda.synchronize()
    r0, r1 = 4000, 4064
    a = A[r0:r1].cpu().numpy()
    b = Bm.cpu().numpy()
    ref = orc.gemm(a, b, C0[r0:r1].cpu().numpy(), 1.0)
    got = C[r0:r1].cpu().numpy()
    # |err| <= 1e-12 * (|a| |b|)_ij: a bound proportional to the entry's own magnitude sum
    bound = np.abs(a) @ np.abs(b) + np.abs(C0[r0:r1].cpu().numpy())
    assert np.all(np.abs(got - ref) <= 1e-12 * bound)
    # Freivalds: (C - C0) x == A (B x) for random x, to the rounding of |A||B||x|
    x = torch.randn((n, 4), generator=g, dtype=torch.float64, device="cuda")
    lhs = (C - C0) @ x
    rhs = A @ (Bm @ x)
    scale = A.abs() @ (Bm.abs() @ x.abs())
    assert float(((lhs - rhs).abs() / scale).max().item()) < 1e-12


@pytest.mark.parametrize("upper", [False, True])
def test_dsyrk_8192(handle, upper):
    """C = -E E^T + C on one triangle, n = k = 8192, E N(0,1): a 64-row slab
    against the oracle's kernels::gemm (SYRK's entries are those of E E^T), the
    other triangle bit for bit, and a Freivalds probe of the symmetric
    completion."""
    n = 8192
    _say(f"dsyrk n=k={n} upper={upper}")
    g = torch.Generator(device="cuda").manual_seed(82)
    E = la.colmajor(torch.randn((n, n), generator=g, dtype=torch.float64, device="cuda"))
    C0 = la.colmajor(torch.randn((n, n), generator=g, dtype=torch.float64, device="cuda"))
    C = la.colmajor(C0.clone())
    la.syrk(handle, E, C, alpha=-1.0, beta=1.0, upper=upper)
    torch.cuda.synchronize()
    if upper:
        assert torch.equal(torch.tril(C, -1), torch.tril(C0, -1))
    else:
        assert torch.equal(torch.triu(C, 1), torch.triu(C0, 1))
    r0, r1 = 6000, 6064
    e = E.cpu().numpy()
    c0 = C0[r0:r1].cpu().numpy()
    ref = orc.gemm(e[r0:r1], e.T.copy(), c0, -1.0)
    got = C[r0:r1].cpu().numpy()
    cols = np.arange(n)[None, :]
    rows = np.arange(r0, r1)[:, None]
    mask = (cols >= rows) if upper else (cols <= rows)
    bound = np.abs(e[r0:r1]) @ np.abs(e.T) + np.abs(c0)
    assert np.all(np.abs(got - ref)[mask] <= 1e-12 * bound[mask])
    # Freivalds on the symmetric completion S of the updated triangle
    D = C - C0
    S = torch.triu(D) + torch.triu(D, 1).t() if upper else torch.tril(D) + torch.tril(D, -1).t()
    x = torch.randn((n, 4), generator=g, dtype=torch.float64, device="cuda")
    lhs = S @ x
    rhs = -(E @ (E.t() @ x))
    scale = E.abs() @ (E.abs().t() @ x.abs())
    assert float(((lhs - rhs).abs() / scale).max().item()) < 1e-12


# ------------------------------------------------------ mixed, n = 16384 --
@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mixed_n16384(handle, prec):
    """The mixed-precision configuration at full size: RBF n = 16384, 64 RHS;
    the backward error <= 1e-14 recomputed in FP64 with torch, and X against
    the FP64 potrf + potrs solution to cond(A) * 1e-15."""
    n, nrhs = 16384, 64
    _say(f"mixed n={n} {prec}")
    A = rbf_gpu(n)
    B = la.colmajor(torch.from_numpy(synth.rhs(n, nrhs, seed=1)).cuda())
    X = la.empty_colmajor(n, nrhs)
    it, berr = la.dsposv(handle, A, B, X, prec=prec, tol=1e-14, max_iter=30)
    torch.cuda.synchronize()
    eta = eta_torch(A, X, B)
    _say(f"  iters {it}  reported {berr:.2e}  recomputed {eta:.2e}")
    assert it >= 1, it  # refinement ran, no fallback to the FP64 factorisation
    assert berr <= 1e-14 and eta <= 1e-14
    L = la.colmajor(A.clone())
    assert int(la.potrf(handle, L).item()) == 0
    X64 = la.colmajor(B.clone())
    la.potrs(handle, L, X64)
    cond_bound = float(A.abs().sum(dim=1).max().item()) / 1e-2  # lambda_min >= 1e-2
    rel = float(((X - X64).abs().amax(dim=0) / X64.abs().amax(dim=0)).max().item())
    _say(f"  |X - X64|/|X64| {rel:.2e}  cond bound {cond_bound:.2e}")
    assert rel <= cond_bound * 1e-15


# ----------------------------------------------------------- batched 4096 --
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batched_full_config(handle, dtype):
    """4096 RBF systems, n_i uniform in [16, 512] (own points n_i x 4 per
    system, lengthscale 0.5, + 1e-3 I), 8 RHS, in one launch: every system's
    info is 0 and its backward error (FP64, torch) is at the precision's level."""
    _say(f"batched 4096 {dtype}")
    sizes = [int(n) for n in synth.batched_sizes(4096, 16, 512, seed=2)]
    nrhs = 8
    A0, B0, As, Bs = [], [], [], []
    for i, n in enumerate(sizes):
        x = torch.from_numpy(synth.points(n, 4, seed=2, stream=100 + i)).cuda()
        sq = (x * x).sum(dim=1)
        d2 = (sq[:, None] + sq[None, :] - 2.0 * (x @ x.t())).clamp_min_(0.0)
        a = torch.exp(d2 * (-1.0 / (2.0 * 0.5 * 0.5)))
        a = 0.5 * (a + a.t())
        a.diagonal().add_(1e-3)
        b = torch.from_numpy(synth.rhs(n, nrhs, seed=3, stream=200 + i)).cuda()
        A0.append(a)
        B0.append(b)
        As.append(la.colmajor(a.to(dtype)))
        Bs.append(la.colmajor(b.to(dtype)))
    batch = la.Batch(As, Bs, order=la.lpt_order(sizes))
    info = la.posv_batched(handle, batch)
    torch.cuda.synchronize()
    assert int((info != 0).sum().item()) == 0
    worst = 0.0
    for a, b, X in zip(A0, B0, Bs):
        worst = max(worst, eta_torch(a, X.double(), b))
    _say(f"  worst backward error {worst:.2e}")
    assert worst <= (1e-15 if dtype == torch.float64 else 1e-6)


# ----------------------------------------------------------- adjoint 8192 --
def test_adjoint_n8192_closed_form_and_32_directions(handle):
    """GP negative log marginal likelihood at n = 8192 (RBF of points in
    [0,1]^8, lengthscale 0.3, + 1e-2 I; y N(0,1)).  The gradient from
    rmh_dpotrf_adjoint equals 0.5 (A^-1 - alpha alpha^T) computed with
    torch.linalg (rocSOLVER), and matches central differences with h = 1e-6
    on 32 random symmetric directions."""
    n = 8192
    _say("adjoint n=8192")
    A = rbf_gpu(n)
    y = torch.from_numpy(synth.normal(11, 0, n)).cuda()

    def nlml_and_factor(M):
        L = la.colmajor(M.clone())
        assert int(la.potrf(handle, L).item()) == 0
        Lt = torch.tril(L)
        z = torch.linalg.solve_triangular(Lt, y[:, None], upper=False)[:, 0]
        return 0.5 * torch.dot(z, z).item() + torch.log(torch.diagonal(Lt)).sum().item(), Lt, z

    f0, Lt, z = nlml_and_factor(A)
    w = torch.linalg.solve_triangular(Lt.t(), z[:, None], upper=True)
    Lbar = la.colmajor(torch.tril(-w @ z[None, :]))
    Lbar.diagonal().add_(1.0 / Lt.diagonal())
    G = la.potrf_adjoint(handle, la.colmajor(Lt), Lbar)
    torch.cuda.synchronize()
    # closed form with torch.linalg, independent of this library
    Ainv = torch.cholesky_inverse(torch.linalg.cholesky(A))
    alpha = Ainv @ y
    ref = 0.5 * (Ainv - torch.outer(alpha, alpha))
    ref = 0.5 * (ref + ref.t())
    err = float(((G - ref).abs().max() / ref.abs().max()).item())
    _say(f"  closed form rel err {err:.2e}")
    assert err <= 1e-8
    assert torch.equal(G, G.t())
    del Ainv, ref

    def nlml_torch(M):
        Lt = torch.linalg.cholesky(M)
        z = torch.linalg.solve_triangular(Lt, y[:, None], upper=False)[:, 0]
        return 0.5 * torch.dot(z, z).item() + torch.log(torch.diagonal(Lt)).sum().item()

    # Central differences at h = 1e-6 along N(0,1) directions carry two
    # errors of their own: rounding (the two evaluations' errors over 2h),
    # estimated per direction from a second, independent evaluation with
    # torch.linalg.cholesky; and truncation, O(h^2) because h ||A^-1 V|| is
    # about 1e-2 here (lambda_min(A) = 1e-2), estimated from the quotient at
    # h/2.  The mismatch must stay within those estimates plus 1e-5 |an|, the
    # Richardson value (4 D(h/2) - D(h)) / 3 within 1e-5 |an| plus the
    # rounding, and both below 1e-2 relative in any case (a wrong adjoint,
    # such as a missing diagonal halving, is off by O(1)).
    gen = torch.Generator(device="cuda").manual_seed(1234)
    h = 1e-6
    worst = worst_rich = 0.0
    for _ in range(32):
        V = torch.randn((n, n), generator=gen, dtype=torch.float64, device="cuda")
        V = 0.5 * (V + V.t())
        quot, noise = [], 0.0
        for hh in (h, h / 2):
            Ap, Am = A + hh * V, A - hh * V
            fp, _, _ = nlml_and_factor(Ap)
            fm, _, _ = nlml_and_factor(Am)
            noise = max(noise, (abs(fp - nlml_torch(Ap)) + abs(fm - nlml_torch(Am))) / (2 * hh))
            quot.append((fp - fm) / (2 * hh))
            del Ap, Am
        d1, d2 = quot
        an = float((G * V).sum().item())
        trunc = 4.0 / 3.0 * abs(d1 - d2)
        rich = (4 * d2 - d1) / 3
        assert abs(d1 - an) <= 1e-5 * abs(an) + 2 * trunc + 4 * noise, (d1, d2, an, noise)
        assert abs(rich - an) <= 1e-5 * abs(an) + 8 * noise, (d1, d2, an, noise)
        assert abs(d1 - an) <= 1e-2 * max(1.0, abs(an)), (d1, an)
        worst = max(worst, abs(d1 - an) / max(1.0, abs(an)))
        worst_rich = max(worst_rich, abs(rich - an) / max(1.0, abs(an)))
        del V
    _say(f"  Richardson mismatch {worst_rich:.2e} relative")
    _say(f"  worst central-difference mismatch {worst:.2e} relative  (f = {f0:.6e})")


def test_adjoint_n2048_vs_cpu_reference(handle):
    """The device adjoint against the CPU FP64 reference adjoint at n = 2048."""
    from oracle import adjoint as adj
    n = 2048
    _say("adjoint n=2048 vs CPU reference")
    A = synth.gram_spd(n, seed=5)
    L = np.linalg.cholesky(A)
    rng = np.random.default_rng(5)
    Lbar = np.tril(rng.standard_normal((n, n)))
    ref = adj.potrf_adjoint(L, Lbar)
    got = la.potrf_adjoint(handle, dev(L), dev(Lbar))
    G = host(got)
    np.testing.assert_allclose(G, ref, rtol=0, atol=1e-10 * np.max(np.abs(ref)))
