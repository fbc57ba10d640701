"""Batched uneven SPD factor + solve (rmh_?posv_batched and friends).

No reference counterpart: every system must equal the single-matrix path --
the LAPACK / oracle factor to rounding, backward error of the solve at FP64
level, LAPACK info on failure -- and the launch must replay from a captured
graph.  Inputs follow the batched workload: RBF matrices of points in
[0,1]^4, lengthscale 0.5, + 1e-3 I, sizes drawn uniformly from [16, 512].
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from remora_amd import capi, synth
from remora_amd import linalg as la

pytestmark = pytest.mark.gpu


def rbf_batch(sizes, seed=2):
    mats = []
    for i, n in enumerate(sizes):
        x = synth.points(int(n), 4, seed=seed, stream=100 + i)
        mats.append(synth.rbf(x, 0.5, 1e-3))
    return mats


def rhs_batch(sizes, nrhs, seed=3):
    return [synth.rhs(int(n), nrhs, seed=seed, stream=200 + i) for i, n in enumerate(sizes)]


def to_dev(mats, dtype):
    return [la.colmajor(torch.from_numpy(np.asarray(m, dtype=np.float64)).to(dtype).cuda()) for m in mats]


SIZES = [1, 2, 15, 16, 17, 31, 32, 33, 63, 64, 65, 100, 255, 256, 257, 511, 512] + list(
    synth.batched_sizes(31, 16, 512, seed=2))


@pytest.fixture(scope="module")
def h():
    return capi.Handle(0)


def backward_error(A, X, B):
    return synth.backward_error(A, X, B)


@pytest.mark.parametrize("nb", [16, 32])
def test_dposv_batched_matches_lapack(h, nb):
    A = rbf_batch(SIZES)
    B = rhs_batch(SIZES, 8)
    dA, dB = to_dev(A, torch.float64), to_dev(B, torch.float64)
    batch = la.Batch(dA, dB)
    info = la.posv_batched(h, batch, cfg=capi.BatchConfig(nb, 0, 0))
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [0] * len(SIZES)
    for i, n in enumerate(SIZES):
        F = dA[i].cpu().numpy()
        L = np.tril(F)
        Lref = np.linalg.cholesky(A[i])
        # the factor: LAPACK's to rounding (forward error ~ cond * u)
        assert np.max(np.abs(L - Lref)) <= 1e-9 * np.max(np.abs(Lref)), n
        assert np.linalg.norm(A[i] - L @ L.T) <= 1e-14 * n * np.linalg.norm(A[i]), n
        # the strict upper triangle is untouched
        iu = np.triu_indices(n, 1)
        assert np.array_equal(F[iu], A[i][iu]), n
        X = dB[i].cpu().numpy()
        assert backward_error(A[i], X, B[i]) <= 1e-15 * max(8, n ** 0.5), n


def test_dposv_batched_vs_oracle(h):
    sizes = [40, 200, 512]
    A = rbf_batch(sizes, seed=7)
    B = rhs_batch(sizes, 4, seed=8)
    dA, dB = to_dev(A, torch.float64), to_dev(B, torch.float64)
    la.posv_batched(h, la.Batch(dA, dB))
    torch.cuda.synchronize()
    for i, n in enumerate(sizes):
        Xo, Lo, info, _ = orc.posv(A[i], B[i])
        assert info == 0
        L = np.tril(dA[i].cpu().numpy())
        np.testing.assert_allclose(L, np.tril(Lo), rtol=0, atol=1e-10 * np.max(np.abs(Lo)))
        np.testing.assert_allclose(dB[i].cpu().numpy(), Xo, rtol=1e-8, atol=1e-8 * np.max(np.abs(Xo)))


def test_sposv_batched(h):
    A = rbf_batch(SIZES)
    B = rhs_batch(SIZES, 8)
    dA, dB = to_dev(A, torch.float32), to_dev(B, torch.float32)
    info = la.posv_batched(h, la.Batch(dA, dB))
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [0] * len(SIZES)  # cond <= 2e5 here: FP32 succeeds
    for i, n in enumerate(SIZES):
        L = np.tril(dA[i].cpu().numpy().astype(np.float64))
        assert np.linalg.norm(A[i] - L @ L.T) <= 1e-6 * np.linalg.norm(A[i]), n
        X = dB[i].cpu().numpy().astype(np.float64)
        assert backward_error(A[i], X, B[i]) <= 1e-5, n


def test_potrf_then_potrs_equals_posv(h):
    sizes = [17, 100, 300, 512]
    A = rbf_batch(sizes)
    B = rhs_batch(sizes, 5)
    a1, b1 = to_dev(A, torch.float64), to_dev(B, torch.float64)
    a2, b2 = to_dev(A, torch.float64), to_dev(B, torch.float64)
    la.posv_batched(h, la.Batch(a1, b1))
    bt = la.Batch(a2, b2)
    la.potrf_batched(h, bt)
    la.potrs_batched(h, bt)
    torch.cuda.synchronize()
    for i in range(len(sizes)):
        assert torch.equal(a1[i], a2[i]) and torch.equal(b1[i], b2[i])


def test_order_does_not_change_results(h):
    sizes = list(synth.batched_sizes(24, 16, 512, seed=5))
    A = rbf_batch(sizes)
    B = rhs_batch(sizes, 8)
    a1, b1 = to_dev(A, torch.float64), to_dev(B, torch.float64)
    a2, b2 = to_dev(A, torch.float64), to_dev(B, torch.float64)
    la.posv_batched(h, la.Batch(a1, b1))
    order = la.lpt_order(sizes)
    assert list(np.asarray(sizes)[order]) == sorted(sizes, reverse=True)
    la.posv_batched(h, la.Batch(a2, b2, order=order))
    torch.cuda.synchronize()
    for i in range(len(sizes)):
        assert torch.equal(a1[i], a2[i]) and torch.equal(b1[i], b2[i])


def test_info_and_edge_sizes(h):
    n = 70
    bad = np.diag(np.arange(1.0, n + 1))
    bad[44, 44] = -1.0  # leading minor of order 45 fails
    good = rbf_batch([33])[0]
    mats = [bad, good, np.zeros((0, 0))]
    dA = to_dev(mats, torch.float64)
    dB = to_dev([np.ones((n, 2)), np.ones((33, 2)), np.zeros((0, 2))], torch.float64)
    info = la.posv_batched(h, la.Batch(dA, dB))
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [45, 0, 0]
    # the factor before the failing block column is complete
    L = np.tril(dA[0].cpu().numpy())
    np.testing.assert_allclose(np.diag(L)[:32], np.sqrt(np.arange(1.0, 33)), rtol=1e-15)


def test_invalid_arguments(h):
    dA = to_dev([np.eye(4)], torch.float64)
    b = la.Batch(dA)
    lib = capi.load()
    st = lib.rmh_dpotrf_batched(h.value, 1, 600, la._bp(b.n), la._bp(b.pa), la._bp(b.lda), la._bp(b.info), None,
                                None, la._stream(None))
    assert st == capi.INVALID_VALUE
    cfg = capi.BatchConfig(8, 0, 0)
    st = lib.rmh_dpotrf_batched(h.value, 1, 4, la._bp(b.n), la._bp(b.pa), la._bp(b.lda), la._bp(b.info), None,
                                C_byref(cfg), la._stream(None))
    assert st == capi.INVALID_VALUE


def C_byref(x):
    import ctypes
    return ctypes.byref(x)


def test_graph_capture_replays(h):
    sizes = list(synth.batched_sizes(40, 16, 512, seed=11))
    A = rbf_batch(sizes, seed=12)
    B = rhs_batch(sizes, 8, seed=13)
    dA, dB = to_dev(A, torch.float64), to_dev(B, torch.float64)
    A0 = [a.clone() for a in dA]
    B0 = [b.clone() for b in dB]
    batch = la.Batch(dA, dB, order=la.lpt_order(sizes))
    # eager result (also makes the handle's one-time allocation)
    la.posv_batched(h, batch)
    torch.cuda.synchronize()
    ref_a = [a.clone() for a in dA]
    ref_b = [b.clone() for b in dB]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    for a, a0 in zip(dA, A0):
        a.copy_(a0)
    for b, b0 in zip(dB, B0):
        b.copy_(b0)
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        la.posv_batched(h, batch)
    for rep in range(3):
        # fresh inputs by eager copies between replays
        for a, a0 in zip(dA, A0):
            a.copy_(a0)
        for b, b0 in zip(dB, B0):
            b.copy_(b0)
        batch.info.fill_(-7)
        g.replay()
        torch.cuda.synchronize()
        assert batch.info.cpu().tolist() == [0] * len(sizes), rep
        for i in range(len(sizes)):
            assert torch.equal(dA[i], ref_a[i]) and torch.equal(dB[i], ref_b[i]), (rep, i)


def test_autotune(h):
    sizes = list(synth.batched_sizes(64, 16, 512, seed=2))
    cfg, ms = la.batched_autotune(h, sizes, 8, fp64=True)
    assert cfg.nb in (16, 32) and cfg.wg_per_cu >= 1 and ms > 0
    # the tuned configuration is used when none is given
    A = rbf_batch(sizes[:8])
    dA = to_dev(A, torch.float64)
    info = la.potrf_batched(h, la.Batch(dA))
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [0] * 8
