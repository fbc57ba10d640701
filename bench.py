"""Benchmark: potrf+solve TFLOP/s at n=16384, FP64 vs FP32-MFMA + FP64 refinement.

A step is Remora's solve(A, B, symm_pos_def(), left()) on device-resident
inputs with 64 right-hand sides:
  fp64   copy A into the factor buffer and B into X (the copies
         cholesky_decomposition and matrix_matrix_solve make), dpotrf, dpotrs;
  mixed  rmh_dsposv: FP32 copy of A, Cholesky with FP32-MFMA trailing updates
         and FP64 diagonal blocks, FP32 solve, then FP64 residual + FP32
         correction until the backward error is <= 1e-14 (every iteration is
         inside the timed step).
Work per step is the algorithmic n^3/3 + 2 n^2 nrhs flops for both.

Workloads (both at n=16384, 64 RHS, B U[-1,1]):
  rbf   A_ij = exp(-|x_i - x_j|^2 / (2 * 0.3^2)) + 1e-2 delta_ij, x_i in
        [0,1]^8 (project RNG, seed 1): the mixed-precision comparison, both
        paths on the same matrix;
  gram  A = G G^T / n + 0.1 I with G N(0,1) (seeded torch generator): the
        plain FP64 drop-in case.
`value` is the mixed path's throughput on rbf at backward error <= 1e-14;
the FP64 throughputs on rbf and gram and the ratio stand beside it.  Every
result carries ||A - L L^T||_F/||A||_F (sampled columns) and the backward
error max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf).

The roofline objects are measured live: HIP events recorded by the library
on the launching stream around every trailing-update launch (the dominant
kernel) during the timed steps.  The CPU baseline is the oracle (a
restatement of Remora's default CPU kernels) on one core, on a bounded
sample: potrf + potrs at n=8192 with 64 RHS (about 15 s of CPU time).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64": 78.6e12, "f32": 157.3e12}
HBM_PEAK = 8.0e12
TOL = 1e-14


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_gram(n, nrhs, seed=0):
    import torch
    from remora_amd import linalg as la
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    G = torch.randn((n, n), generator=g, dtype=torch.float64, device="cuda")
    A = torch.empty((n, n), dtype=torch.float64, device="cuda")
    A.copy_(G @ G.t())
    del G
    A.div_(n)
    A = 0.5 * (A + A.t())
    A.diagonal().add_(0.1)
    A = la.colmajor(A)
    B = la.colmajor(torch.rand((n, nrhs), generator=g, dtype=torch.float64, device="cuda") * 2 - 1)
    return A, B


def make_rbf(n, nrhs, d=8, lengthscale=0.3, jitter=1e-2, seed=1):
    import torch
    from remora_amd import linalg as la
    from remora_amd import synth
    x = torch.from_numpy(synth.points(n, d, seed)).cuda()
    sq = (x * x).sum(dim=1)
    d2 = (sq[:, None] + sq[None, :] - 2.0 * (x @ x.t())).clamp_min_(0.0)
    A = torch.exp(d2.mul_(-1.0 / (2.0 * lengthscale * lengthscale)))
    del d2
    A = 0.5 * (A + A.t())
    A.diagonal().add_(jitter)
    A = la.colmajor(A)
    B = la.colmajor(torch.from_numpy(synth.rhs(n, nrhs, seed=seed)).cuda())
    return A, B


def backward_error(h, A, X, B):
    """max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf) on the GPU."""
    from remora_amd import linalg as la
    R = la.colmajor(B.clone())
    la.gemm(h, A, X, R, alpha=-1.0, beta=1.0)
    anorm = A.abs().sum(dim=1).max()
    num = R.abs().amax(dim=0)
    den = anorm * X.abs().amax(dim=0) + B.abs().amax(dim=0)
    return float((num / den).max().item())


def factor_error_sampled(A, L, cols=64, seed=1):
    """||A - L L^T||_F / ||A||_F estimated on sampled columns of A."""
    import torch
    n = A.shape[0]
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    idx = torch.randperm(n, generator=g)[:cols].cuda()
    Lt = torch.tril(L.to(torch.float64))
    Ac = A[:, idx]
    Rc = Ac - Lt @ Lt[idx, :].t()
    return float((Rc.norm() / Ac.norm()).item())


def timed(step, steps, warmup, h, profile):
    """Run warm-up steps, then time `steps` steps with events; returns per-step ms and the trailing profile."""
    import torch
    from remora_amd import linalg as la
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    prof = la.Profile(h)
    if profile:
        prof.reset()
        prof.enable(("trailing_update",))
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    t0 = time.perf_counter()
    ev[0].record()
    for i in range(steps):
        step()
        ev[i + 1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [ev[i].elapsed_time(ev[i + 1]) for i in range(steps)]
    tr = None
    if profile:
        prof.disable()
        tr = prof.read("trailing_update")
    return per, wall, tr


def run_fp64(h, A, B, steps, warmup, profile=True):
    import torch
    from remora_amd import linalg as la
    n, nrhs = B.shape
    L = la.empty_colmajor(n, n)
    X = la.empty_colmajor(n, nrhs)
    info = torch.zeros(1, dtype=torch.int64, device="cuda")

    def step():
        L.copy_(A)
        X.copy_(B)
        la.potrf(h, L, info=info)
        la.potrs(h, L, X)

    per, wall, tr = timed(step, steps, warmup, h, profile)
    assert int(info.item()) == 0, f"potrf info {int(info.item())}"
    return {"wall_s": wall, "step_ms": per, "trailing": tr, "L": L, "X": X}


def run_mixed(h, A, B, steps, warmup, prec="fp32", profile=True):
    from remora_amd import linalg as la
    n, nrhs = B.shape
    X = la.empty_colmajor(n, nrhs)
    hist = []

    def step():
        hist.append(la.dsposv(h, A, B, X, prec=prec, tol=TOL, max_iter=30))

    per, wall, tr = timed(step, steps, warmup, h, profile)
    return {"wall_s": wall, "step_ms": per, "trailing": tr, "X": X, "iters": [i for i, _ in hist[warmup:]],
            "berr": [b for _, b in hist[warmup:]]}


def make_batched(count=4096, lo=16, hi=512, nrhs=8, dtype=None, seed=2):
    """4096 RBF systems: n_i uniform in [16, 512] (project RNG, seed 2), points
    n_i x 4 in [0,1] (own stream per system), lengthscale 0.5, + 1e-3 I; B
    U[-1,1] with 8 columns.  Built on the GPU into two packed buffers."""
    import torch
    from remora_amd import synth
    dtype = dtype or torch.float64
    sizes = [int(n) for n in synth.batched_sizes(count, lo, hi, seed=seed)]
    offs, boffs, tot, btot = [], [], 0, 0
    for n in sizes:
        offs.append(tot)
        boffs.append(btot)
        tot += n * n
        btot += n * nrhs
    A0 = torch.empty(tot, dtype=torch.float64, device="cuda")
    B0 = torch.empty(btot, dtype=torch.float64, device="cuda")
    for i, n in enumerate(sizes):
        x = torch.from_numpy(synth.points(n, 4, seed=seed, stream=100 + i)).cuda()
        sq = (x * x).sum(dim=1)
        d2 = (sq[:, None] + sq[None, :] - 2.0 * (x @ x.t())).clamp_min_(0.0)
        a = torch.exp(d2 * (-1.0 / (2.0 * 0.5 * 0.5)))
        a = 0.5 * (a + a.t())
        a.diagonal().add_(1e-3)
        A0[offs[i]:offs[i] + n * n].copy_(a.t().reshape(-1))  # column-major
        B0[boffs[i]:boffs[i] + n * nrhs].copy_(
            torch.from_numpy(synth.rhs(n, nrhs, seed=seed + 1, stream=200 + i)).cuda().t().reshape(-1))
    return sizes, offs, boffs, A0.to(dtype), B0.to(dtype)


def run_batched(h, steps, warmup, dtype_name):
    """Autotune, then time rmh_?posv_batched (factor + 8-RHS solve of all
    4096 systems in one persistent launch, replayed from a captured graph)."""
    import torch
    from remora_amd import linalg as la
    dtype = torch.float64 if dtype_name == "f64" else torch.float32
    nrhs = 8
    sizes, offs, boffs, A0, B0 = make_batched(nrhs=nrhs, dtype=dtype)
    A, B = A0.clone(), B0.clone()
    As = [A[o:o + n * n].view(n, n).t() for o, n in zip(offs, sizes)]
    Bs = [B[o:o + n * nrhs].view(nrhs, n).t() for o, n in zip(boffs, sizes)]
    cfg, tune_ms = la.batched_autotune(h, sizes, nrhs, fp64=(dtype_name == "f64"))
    order = la.lpt_order(sizes) if cfg.lpt_order else None
    batch = la.Batch(As, Bs, order=order)
    la.posv_batched(h, batch)  # eager once
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        la.posv_batched(h, batch)
    times = []
    for i in range(warmup + steps):
        A.copy_(A0)
        B.copy_(B0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        g.replay()
        e1.record()
        torch.cuda.synchronize()
        if i >= warmup:
            times.append(e0.elapsed_time(e1))
    info = batch.info.cpu()
    assert int((info != 0).sum()) == 0, "batched: a system failed to factor"
    # accuracy on a sample of systems (host FP64 check)
    import numpy as np
    from remora_amd import synth
    worst_eta, worst_ferr = 0.0, 0.0
    for i in list(range(0, len(sizes), 512)) + [int(np.argmax(sizes))]:
        n = sizes[i]
        a = A0[offs[i]:offs[i] + n * n].double().view(n, n).t().cpu().numpy()
        b = B0[boffs[i]:boffs[i] + n * nrhs].double().view(nrhs, n).t().cpu().numpy()
        L = np.tril(As[i].double().cpu().numpy())
        x = Bs[i].double().cpu().numpy()
        worst_eta = max(worst_eta, synth.backward_error(a, x, b))
        worst_ferr = max(worst_ferr, float(np.linalg.norm(a - L @ L.T) / np.linalg.norm(a)))
    ms = median(times)
    n_arr = np.asarray(sizes, dtype=np.float64)
    work = float(np.sum(n_arr ** 3 / 3 + 2 * n_arr ** 2 * nrhs))
    esize = 8 if dtype_name == "f64" else 4
    nbytes = float(np.sum(2 * n_arr ** 2 + 2 * nrhs * n_arr)) * esize
    achieved = work / (ms * 1e-3)
    peak = PEAK[dtype_name]
    ceiling = min(peak, 8e12 * work / nbytes)
    return {"tflops": achieved / 1e12, "ms": ms, "step_ms": times, "systems": len(sizes), "nrhs": nrhs,
            "work_flops": work, "algorithmic_bytes": nbytes,
            "config": {"nb": cfg.nb, "wg_per_cu": cfg.wg_per_cu, "lpt_order": cfg.lpt_order, "autotune_ms": tune_ms},
            "accuracy": {"max_backward_error_sampled": worst_eta, "max_factor_error_sampled": worst_ferr},
            "roofline": {"bound": "hbm" if ceiling < peak else "mfma", "achieved": achieved / 1e12,
                         "peak": ceiling / 1e12, "unit": "TFLOP/s", "frac": achieved / ceiling,
                         "intensity_flop_per_byte": work / nbytes}}


def run_adjoint(h, steps, warmup, n=8192):
    """GP negative log marginal likelihood at n=8192 (RBF of points in
    [0,1]^8, lengthscale 0.3, + 1e-2 I; y N(0,1)): forward potrf, the factor's
    gradient Lbar = tril(-w z^T) + diag(1/L_ii) (z = L^-1 y, w = L^-T z), then
    the timed step: rmh_dpotrf_adjoint.  Work 7n^3/3 (the L^T Lbar product
    on its triangle, n^3/3, and two n x n triangular solves, n^3 each)."""
    import torch
    from remora_amd import linalg as la
    from remora_amd import synth
    A, _ = make_rbf(n, 1, d=8, lengthscale=0.3, jitter=1e-2, seed=1)
    y = torch.from_numpy(synth.normal(11, 0, n)).cuda()
    L = la.colmajor(A.clone())
    info = la.potrf(h, L)
    assert int(info.item()) == 0
    Lt = torch.tril(L)
    z = torch.linalg.solve_triangular(Lt, y[:, None], upper=False)
    w = torch.linalg.solve_triangular(Lt.t(), z, upper=True)
    Lbar = la.colmajor(torch.tril(-w @ z.t()))
    Lbar.diagonal().add_(1.0 / Lt.diagonal())
    Abar = la.empty_colmajor(n, n)
    times = []
    for i in range(warmup + steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        la.potrf_adjoint(h, L, Lbar, Abar)
        e1.record()
        torch.cuda.synchronize()
        if i >= warmup:
            times.append(e0.elapsed_time(e1))
    # closed form 0.5 (A^-1 - alpha alpha^T) with the library's own solve
    Ainv = la.colmajor(torch.eye(n, dtype=torch.float64, device="cuda"))
    la.potrs(h, L, Ainv)
    alpha = Ainv @ y
    ref = 0.5 * (Ainv - torch.outer(alpha, alpha))
    ref = 0.5 * (ref + ref.t())
    err = float(((Abar - ref).abs().max() / ref.abs().max()).item())
    ms = median(times)
    work = 7.0 * n ** 3 / 3.0
    achieved = work / (ms * 1e-3)
    return {"tflops": achieved / 1e12, "ms": ms, "step_ms": times, "n": n, "work_flops": work,
            "accuracy": {"max_rel_err_vs_closed_form": err},
            "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK["f64"] / 1e12, "unit": "TFLOP/s",
                         "frac": achieved / PEAK["f64"]}}


def run_update(h, steps, warmup, n=16384):
    """Rank-one update of a Cholesky factor (cholesky_decomposition::update)
    at n=16384, FP64, lower storage: L tril(N(0,1))/sqrt(n) + 2 I (seed 21),
    v N(0,1)/sqrt(n), alpha 1, beta 0.5.  The factor is restored from a copy
    before each step (outside the timed events).  Algorithmic traffic: the
    stored triangle read once and written once, n(n+1) * 8 bytes (+ v)."""
    import torch
    from remora_amd import linalg as la
    g = torch.Generator(device="cuda").manual_seed(21)
    L0 = la.colmajor(torch.tril(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)) / n ** 0.5
                     + 2.0 * torch.eye(n, dtype=torch.float64, device="cuda"))
    v = torch.randn(n, dtype=torch.float64, device="cuda", generator=g) / n ** 0.5
    L = la.empty_colmajor(n, n)
    info = torch.zeros(1, dtype=torch.int64, device="cuda")
    times = []
    for i in range(warmup + steps):
        L.copy_(L0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        la.chol_update(h, L, 1.0, 0.5, v, info=info)
        e1.record()
        torch.cuda.synchronize()
        assert int(info.item()) == 0
        if i >= warmup:
            times.append(e0.elapsed_time(e1))
    # sampled invariant: (L' L'^T)_ij = (L L^T)_ij + 0.5 v_i v_j
    idx = torch.randint(0, n, (2, 256), generator=g, device="cuda")
    Ln, Lo = torch.tril(L), torch.tril(L0)
    lhs = (Ln[idx[0]] * Ln[idx[1]]).sum(1)
    rhs = (Lo[idx[0]] * Lo[idx[1]]).sum(1) + 0.5 * v[idx[0]] * v[idx[1]]
    err = float(((lhs - rhs).abs().max() / rhs.abs().max()).item())
    ms = median(times)
    nbytes = float(n * (n + 1) * 8 + 2 * n * 8)
    achieved = nbytes / (ms * 1e-3)
    return {"gbps": achieved / 1e9, "ms": ms, "step_ms": times, "n": n, "algorithmic_bytes": nbytes,
            "accuracy": {"max_rel_err_sampled_invariant": err},
            "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK}}


def run_pstrf(h, steps, warmup, n=4096):
    """Pivoted Cholesky (kernels::pstrf<lower>) at n=4096, FP64: the Gram
    matrix G G^T / n + 0.1 I of G N(0,1) (seed 31), full rank.  The matrix is
    restored from a copy before each step (outside the timed events).  Work
    n^3/3 flops, as for potrf (the pivoting is O(n^2))."""
    import torch
    from remora_amd import linalg as la
    g = torch.Generator(device="cuda").manual_seed(31)
    G = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    A0 = la.colmajor(G @ G.t() / n + 0.1 * torch.eye(n, dtype=torch.float64, device="cuda"))
    A = la.empty_colmajor(n, n)
    times = []
    for i in range(warmup + steps):
        A.copy_(A0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        piv, rank = la.pstrf(h, A)
        e1.record()
        torch.cuda.synchronize()
        if i >= warmup:
            times.append(e0.elapsed_time(e1))
    assert int(rank.item()) == n
    # reconstruction on sampled entries: (L L^T)_ij = (P^T A P)_ij
    perm = torch.arange(n, device="cuda")
    for j, p in enumerate(piv.cpu().tolist()):
        if p != j:
            perm[[j, p]] = perm[[p, j]]
    L = torch.tril(A)
    idx = torch.randint(0, n, (2, 256), generator=g, device="cuda")
    lhs = (L[idx[0]] * L[idx[1]]).sum(1)
    rhs = A0[perm[idx[0]], perm[idx[1]]]
    err = float(((lhs - rhs).abs().max() / A0.abs().max()).item())
    ms = median(times)
    work = n ** 3 / 3.0
    achieved = work / (ms * 1e-3)
    return {"tflops": achieved / 1e12, "ms": ms, "step_ms": times, "n": n, "work_flops": work,
            "accuracy": {"max_rel_err_sampled_reconstruction": err},
            "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK["f64"] / 1e12, "unit": "TFLOP/s",
                         "frac": achieved / PEAK["f64"]}}


def run_blas3(h, steps, warmup, n=8192):
    """The standalone FP64 products of the drop-in configuration: GEMM
    C += A B (m = n = k = 8192, the reference's noalias(C) += prod(A, B)) and
    SYRK C += A A^T on the lower triangle (n = k = 8192), with A, B, C
    N(0,1) (seed 3).  Work 2 n^3 and n^2 (n + 1) flops (the usual count; the
    reference's benchmark programs report half of it).  Accuracy: one beta = 0
    product, 64 sampled columns against torch's FP64 product, max relative to
    the column's max |entry|."""
    import torch
    from remora_amd import linalg as la
    g = torch.Generator(device="cuda").manual_seed(3)
    A = la.colmajor(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
    B = la.colmajor(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
    C = la.colmajor(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
    idx = torch.randperm(n, generator=torch.Generator().manual_seed(4))[:64].cuda()
    out = {}

    def rel(got, ref):
        return float(((got - ref).abs().amax(dim=0) / ref.abs().amax(dim=0)).max().item())

    la.gemm(h, A, B, C, alpha=1.0, beta=0.0)
    err_g = rel(C[:, idx], A @ B[:, idx])
    per, _, _ = timed(lambda: la.gemm(h, A, B, C, alpha=1.0, beta=1.0), steps, warmup, h, False)
    ms = median(per)
    out["gemm_f64"] = {"m": n, "n": n, "k": n, "ms": ms, "step_ms": per, "tflops": 2.0 * n ** 3 / (ms * 1e-3) / 1e12,
                       "frac_of_peak": 2.0 * n ** 3 / (ms * 1e-3) / PEAK["f64"], "max_rel_err_sampled": err_g}
    la.syrk(h, A, C, alpha=1.0, beta=0.0)
    ref = A @ A[idx].t()
    mask = torch.arange(n, device="cuda")[:, None] >= idx[None, :]  # the stored (lower) part of each column
    got = C[:, idx]
    err_s = float((((got - ref).abs() * mask).amax(dim=0) / (ref.abs() * mask).amax(dim=0)).max().item())
    per, _, _ = timed(lambda: la.syrk(h, A, C, alpha=1.0, beta=1.0), steps, warmup, h, False)
    ms = median(per)
    work = float(n) * n * (n + 1)
    out["syrk_f64"] = {"n": n, "k": n, "ms": ms, "step_ms": per, "tflops": work / (ms * 1e-3) / 1e12,
                       "frac_of_peak": work / (ms * 1e-3) / PEAK["f64"], "max_rel_err_sampled": err_s}
    return out


def median(v):
    s = sorted(v)
    return s[len(s) // 2]


def measured_traffic(dtype):
    """HBM bytes per launch of the trailing-update kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (calibrated on a device copy of
    known size); None when the profile is absent."""
    path = os.path.join(ROOT, "profiles", "trailing_update_traffic.json")
    try:
        with open(path) as f:
            prof = json.load(f)
        return prof["kernels"]["trail_" + dtype]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def roofline(tr, dtype, kernel):
    achieved = tr["flops"] / (tr["ms"] * 1e-3) if tr and tr["ms"] > 0 else None
    return {"bound": "mfma", "achieved": achieved / 1e12 if achieved else None, "peak": PEAK[dtype] / 1e12,
            "unit": "TFLOP/s", "frac": achieved / PEAK[dtype] if achieved else None,
            "traffic": measured_traffic(dtype),
            "traffic_source": "profiles/trailing_update_traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, "
                              "calibrated on a 2 GiB device copy)",
            "kernel": kernel, "launches": tr["launches"] if tr else None,
            "avg_launch_us": tr["ms"] * 1e3 / tr["launches"] if tr and tr["launches"] else None}


def cpu_baseline(n=8192, nrhs=64, reps=3):
    """Oracle (Remora's default CPU kernels restated) on one core, pinned to
    one CPU like the reference's single-threaded kernels: potrf + potrs of the
    Gram matrix G G^T / n + 0.1 I with 64 RHS at n = 8192 (the value, about
    18 s) and n = 4096, and the CPU configuration itself (n = 512, G G^T / 512
    + 0.1 I, 16 RHS, minimum of 20 runs as examples/Benchmarks/potrf.cpp
    takes)."""
    from oracle import oracle as orc
    from remora_amd import synth
    # inputs first (their numpy products use every core), then one pinned core
    cases = [(n, nrhs, reps, None), (4096, 64, 3, None), (512, 16, 20, 512)]
    inputs = [(synth.gram_spd(m, seed=0, divisor=d), synth.rhs(m, k, seed=0)) for m, k, _, d in cases]
    try:
        old_aff = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {min(old_aff)})
    except (AttributeError, OSError):
        old_aff = None

    def run(A, B, reps):
        m, k = B.shape
        best, eta = float("inf"), None
        for _ in range(reps):
            t0 = time.perf_counter()
            X, L, info, _ = orc.posv(A, B)
            best = min(best, time.perf_counter() - t0)
            assert info == 0
            eta = synth.backward_error(A, X, B)
        work = m ** 3 / 3 + 2 * m * m * k
        return {"n": m, "nrhs": k, "reps": reps, "seconds": best, "tflops": work / best / 1e12,
                "backward_error": eta}

    try:
        res = [run(A, B, c[2]) for (A, B), c in zip(inputs, cases)]
    finally:
        if old_aff is not None:
            os.sched_setaffinity(0, old_aff)
    main, sizes = res[0], res[1:]
    return {"value": main["tflops"], "unit": "TFLOP/s", "cores": 1, "kind": "port",
            "sample": f"oracle potrf+potrs n={n} nrhs={nrhs} (Gram SPD), min of {reps} runs, one thread pinned to "
                      f"one CPU; backward error {main['backward_error']:.2e}; host has {os.cpu_count()} logical CPUs",
            "seconds": main["seconds"], "sizes": sizes}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-rbf", action="store_true", help="skip the RBF legs (FP64 and mixed); profiling runs only")
    ap.add_argument("--no-gram", action="store_true", help="skip the FP64 Gram-matrix leg")
    ap.add_argument("--only-fp64", action="store_true", help="FP64 legs only (no mixed leg)")
    ap.add_argument("--no-p64", action="store_true",
                    help="skip the drop-in legs at n = 4096 and 32768 and the standalone GEMM / SYRK")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-systems legs")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the Cholesky-adjoint leg")
    ap.add_argument("--no-update", action="store_true", help="skip the rank-one update leg")
    ap.add_argument("--no-pstrf", action="store_true", help="skip the pivoted Cholesky leg")
    ap.add_argument("--nb", type=int, default=0, help="potrf panel width (0: library default)")
    ap.add_argument("--nbi", type=int, default=0, help="potrf diagonal block (0: library default)")
    ap.add_argument("--no-lookahead", action="store_true", help="single-stream factorisation")
    ap.add_argument("--lookahead-mode", type=int, default=-1, help="look-ahead mode 0/1/2 (-1: library default)")
    ap.add_argument("--reserve", type=int, default=0,
                    help="compute units per XCD reserved for the panel launches on every look-ahead step (0: none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(the multi-GPU layout is a design, not an implementation)"}))
        sys.exit(2)

    import torch
    from remora_amd import capi
    from remora_amd import linalg as la
    n, nrhs = args.n, args.nrhs
    work = n ** 3 / 3 + 2 * n * n * nrhs
    h = capi.Handle(0)
    if args.nb or args.nbi:
        la.set_potrf_blocking(h, args.nb, args.nbi)
    if args.no_lookahead:
        la.set_potrf_lookahead(h, False)
    if args.lookahead_mode >= 0:
        la.set_potrf_lookahead(h, args.lookahead_mode)
    if args.reserve > 0:
        la.set_potrf_reserve(h, args.reserve)
    paths, acc = {}, {}

    roof64 = roof32 = None
    if not args.no_rbf:
        log(f"[bench] rbf inputs n={n} nrhs={nrhs}")
        A, B = make_rbf(n, nrhs)
        torch.cuda.synchronize()
        log("[bench] rbf fp64 potrf+potrs")
        r = run_fp64(h, A, B, args.steps, args.warmup)
        med64 = median(r["step_ms"])
        eta = backward_error(h, A, r["X"], B)
        ferr = factor_error_sampled(A, r["L"])
        paths["rbf_fp64"] = {"tflops": work / (med64 * 1e-3) / 1e12, "ms": med64, "step_ms": r["step_ms"]}
        acc["rbf_fp64"] = {"backward_error": eta, "factor_error_sampled": ferr}
        roof64 = roofline(r["trailing"], "f64", "gemm_kernel<double,128,128,16,false,true,1,0>: trailing SYRK update, "
                                                      "lower trapezoid, K = panel width (1024 for FP64)")
        log(f"[bench] rbf fp64: median {med64:.3f} ms  {paths['rbf_fp64']['tflops']:.2f} TFLOP/s  eta {eta:.2e}  "
            f"ferr {ferr:.2e}")
        del r
        if not args.only_fp64:
            log("[bench] rbf mixed fp32 dsposv")
            r = run_mixed(h, A, B, args.steps, args.warmup, "fp32")
            medmx = median(r["step_ms"])
            eta = backward_error(h, A, r["X"], B)
            paths["rbf_mixed_fp32"] = {"tflops": work / (medmx * 1e-3) / 1e12, "ms": medmx, "step_ms": r["step_ms"],
                                       "iters": r["iters"]}
            acc["rbf_mixed_fp32"] = {"backward_error": eta, "reported_berr": r["berr"]}
            roof32 = roofline(r["trailing"], "f32", "gemm_kernel<float,128,128,32,false,true,1,0>: trailing SYRK update, "
                                                      "lower trapezoid, K = panel width (1024 for FP32)")
            log(f"[bench] rbf mixed: median {medmx:.3f} ms  {paths['rbf_mixed_fp32']['tflops']:.2f} TFLOP/s  "
                f"eta {eta:.2e}  iters {r['iters']}")
            assert max(r["iters"]) >= 0 and eta <= TOL, "mixed path did not reach the tolerance"
            del r
            log("[bench] rbf mixed split-bf16x3 dsposv")
            r = run_mixed(h, A, B, args.steps, args.warmup, "bf16x3", profile=False)
            medb = median(r["step_ms"])
            eta = backward_error(h, A, r["X"], B)
            paths["rbf_mixed_bf16x3"] = {"tflops": work / (medb * 1e-3) / 1e12, "ms": medb, "step_ms": r["step_ms"],
                                         "iters": r["iters"], "fell_back_to_fp64": any(i < 0 for i in r["iters"])}
            acc["rbf_mixed_bf16x3"] = {"backward_error": eta, "reported_berr": r["berr"]}
            log(f"[bench] rbf mixed bf16x3: median {medb:.3f} ms  {paths['rbf_mixed_bf16x3']['tflops']:.2f} TFLOP/s  "
                f"eta {eta:.2e}  iters {r['iters']}")
            del r
        del A, B
        torch.cuda.empty_cache()

    if not args.no_gram:
        log(f"[bench] gram inputs n={n} nrhs={nrhs}")
        A, B = make_gram(n, nrhs)
        torch.cuda.synchronize()
        r = run_fp64(h, A, B, args.steps, args.warmup)
        med = median(r["step_ms"])
        eta = backward_error(h, A, r["X"], B)
        ferr = factor_error_sampled(A, r["L"])
        paths["gram_fp64"] = {"tflops": work / (med * 1e-3) / 1e12, "ms": med, "step_ms": r["step_ms"]}
        acc["gram_fp64"] = {"backward_error": eta, "factor_error_sampled": ferr}
        log(f"[bench] gram fp64: median {med:.3f} ms  {paths['gram_fp64']['tflops']:.2f} TFLOP/s  eta {eta:.2e}  "
            f"ferr {ferr:.2e}")
        del r, A, B
        torch.cuda.empty_cache()

    if not args.no_p64:
        # the drop-in configuration at its other sizes: potrf + potrs with 64
        # RHS on G G^T / n + 0.1 I, and the standalone GEMM and SYRK
        for n2, st2, wu2 in ((4096, args.steps, args.warmup), (32768, 3, 1)):
            log(f"[bench] gram inputs n={n2} nrhs={nrhs}")
            A, B = make_gram(n2, nrhs)
            torch.cuda.synchronize()
            r = run_fp64(h, A, B, st2, wu2, profile=False)
            med = median(r["step_ms"])
            w2 = n2 ** 3 / 3.0 + 2.0 * n2 * n2 * nrhs
            eta = backward_error(h, A, r["X"], B)
            ferr = factor_error_sampled(A, r["L"])
            paths[f"gram_fp64_n{n2}"] = {"tflops": w2 / (med * 1e-3) / 1e12, "ms": med, "step_ms": r["step_ms"],
                                         "frac_of_peak": w2 / (med * 1e-3) / PEAK["f64"]}
            acc[f"gram_fp64_n{n2}"] = {"backward_error": eta, "factor_error_sampled": ferr}
            log(f"[bench] gram fp64 n={n2}: median {med:.3f} ms  {paths[f'gram_fp64_n{n2}']['tflops']:.2f} TFLOP/s  "
                f"eta {eta:.2e}  ferr {ferr:.2e}")
            del r, A, B
            torch.cuda.empty_cache()
        log("[bench] FP64 GEMM 8192^3 and SYRK n=k=8192")
        for k2, v2 in run_blas3(h, args.steps, args.warmup).items():
            paths[k2] = v2
            log(f"[bench] {k2}: median {v2['ms']:.3f} ms  {v2['tflops']:.2f} TFLOP/s  "
                f"({100 * v2['frac_of_peak']:.1f}% of FP64 peak)  err {v2['max_rel_err_sampled']:.2e}")
        torch.cuda.empty_cache()

    if not args.no_batched:
        for dt in ("f64", "f32"):
            log(f"[bench] batched {dt}: 4096 RBF systems, n in [16, 512], 8 RHS")
            r = run_batched(h, args.steps, args.warmup, dt)
            paths[f"batched_{dt}"] = r
            log(f"[bench] batched {dt}: median {r['ms']:.3f} ms  {r['tflops']:.2f} TFLOP/s  "
                f"({100 * r['roofline']['frac']:.1f}% of {r['roofline']['peak']:.1f})  cfg {r['config']}  "
                f"eta {r['accuracy']['max_backward_error_sampled']:.2e}")
            torch.cuda.empty_cache()

    if not args.no_adjoint:
        log("[bench] adjoint: GP NLML gradient, n=8192")
        r = run_adjoint(h, args.steps, args.warmup)
        paths["adjoint_f64"] = r
        log(f"[bench] adjoint: median {r['ms']:.3f} ms  {r['tflops']:.2f} TFLOP/s  "
            f"err vs closed form {r['accuracy']['max_rel_err_vs_closed_form']:.2e}")
        torch.cuda.empty_cache()

    if not args.no_update:
        log("[bench] rank-one Cholesky update, n=16384")
        r = run_update(h, args.steps, args.warmup)
        paths["chol_update_f64"] = r
        log(f"[bench] update: median {r['ms']:.3f} ms  {r['gbps']:.0f} GB/s "
            f"({100 * r['roofline']['frac']:.1f}% of HBM)  invariant err {r['accuracy']['max_rel_err_sampled_invariant']:.2e}")
        torch.cuda.empty_cache()

    if not args.no_pstrf:
        log("[bench] pivoted Cholesky, n=4096")
        r = run_pstrf(h, args.steps, args.warmup)
        paths["pstrf_f64"] = r
        log(f"[bench] pstrf: median {r['ms']:.3f} ms  {r['tflops']:.2f} TFLOP/s  "
            f"reconstruction err {r['accuracy']['max_rel_err_sampled_reconstruction']:.2e}")
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline")
        cpu = cpu_baseline()
        log(f"[bench] cpu: {cpu['value'] * 1e3:.2f} GFLOP/s")

    mixed = "rbf_mixed_fp32" in paths
    if not mixed and "rbf_fp64" not in paths:  # a profiling run without the headline legs
        print(json.dumps({"paths": paths}), flush=True)
        return
    head = paths["rbf_mixed_fp32"] if mixed else paths["rbf_fp64"]
    out = {
        "metric": "potrf+solve TFLOP/s n=16384: FP64 vs FP32-MFMA+refinement @ equal residual",
        "value": head["tflops"],
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic: RBF kernel matrix of 8-d points U[0,1] (lengthscale 0.3, +1e-2 I) and Gram "
                "G G^T/n + 0.1 I; B U[-1,1]",
        "config": {"workload": (f"SPD solve (potrf + potrs), RBF matrix n={n}, {nrhs} RHS, "
                                + ("FP32-MFMA factor + FP64 refinement to backward error <= 1e-14" if mixed
                                   else "FP64")),
                   "n": n, "nrhs": nrhs, "tol": TOL},
        "speedup_mixed_vs_fp64": (paths["rbf_fp64"]["ms"] / paths["rbf_mixed_fp32"]["ms"]) if mixed else None,
        "paths": paths,
        "accuracy": acc,
        "roofline": roof32 if mixed else roof64,
        "roofline_fp64": roof64,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
