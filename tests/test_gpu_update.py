"""Rank-one Cholesky update on the GPU (rmh_?chol_update) through the C-ABI.

Checked against the oracle's restatement of cholesky_decomposition::update
(decompositions.hpp:132-169) on the same inputs, and through properties that
do not depend on the size: the result is the Cholesky factor of
alpha A + beta v v^T, the other triangle and the padding of the leading
dimension are untouched, a failing downdate reports the oracle's column and
leaves the factor unchanged.  The device sweep sums the running vector in a
different order from the reference's column loop, so FP64 agreement is to
1e-12 relative to the largest entry (FP32: 2e-4).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from remora_amd import capi
from remora_amd import linalg as la
from tests import fixtures
from tests.gpu_util import dev, host

pytestmark = pytest.mark.gpu

_H = {}


def handle():
    if "h" not in _H:
        _H["h"] = capi.Handle(0)
    return _H["h"]


def case(n, seed=0, scale=2.0):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n))
    A = G @ G.T / n + 0.5 * np.eye(n)
    u = rng.standard_normal(n)
    return A, scale * u / max(np.linalg.norm(u), 1e-300)


def stored(L, upper, marker):
    """Factor in the requested storage with `marker` in the other triangle."""
    n = L.shape[0]
    F = L.T.copy() if upper else L.copy()
    other = np.tril_indices(n, -1) if upper else np.triu_indices(n, 1)
    F[other] = marker[other]
    return F


def run_update(F, alpha, beta, v, upper, dtype=torch.float64, ld_pad=3, incv=1):
    n = F.shape[0]
    Fd = dev(F, dtype=dtype, ld=n + ld_pad)
    buf = torch.full((max(n * incv, 1),), float("nan"), dtype=dtype, device="cuda")
    vd = buf[::incv][:n] if n else buf[:0]
    vd.copy_(torch.from_numpy(np.asarray(v)).to(dtype))
    info = torch.full((1,), -7, dtype=torch.int64, device="cuda")
    la.chol_update(handle(), Fd, alpha, beta, vd, upper=upper, info=info)
    torch.cuda.synchronize()
    pad = Fd.as_strided((ld_pad, n), (1, n + ld_pad), Fd.storage_offset() + n) if n and ld_pad else None
    return host(Fd), int(info.item()), pad, host(vd) if n else np.zeros(0)


def oracle_update(L, alpha, beta, v):
    U, info = orc.chol_update(L, alpha, beta, v, row_major=False)
    return np.tril(U), info


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("n", [1, 7, 255, 256, 257, 600, 1100])
@pytest.mark.parametrize("alpha,beta", [(1.0, 1.0), (2.5, 0.3), (0.8, -0.05)])
def test_update_matches_oracle(n, alpha, beta, upper):
    A, v = case(n, seed=n)
    L = np.linalg.cholesky(A)
    marker = np.full((n, n), 3.25)
    got, info, pad, vback = run_update(stored(L, upper, marker), alpha, beta, v, upper)
    ref, rinfo = oracle_update(L, alpha, beta, v)
    assert info == rinfo == 0
    Lg = got.T if upper else got
    err = np.max(np.abs(np.tril(Lg) - ref)) / np.max(np.abs(ref))
    assert err <= 1e-12, (n, alpha, beta, upper, err)
    M = alpha * A + beta * np.outer(v, v)
    Ll = np.tril(Lg)
    assert np.linalg.norm(Ll @ Ll.T - M) <= 1e-13 * np.linalg.norm(M)
    other = np.triu_indices(n, 1)
    assert np.all(Lg[other] == 3.25), "the other triangle changed"
    assert pad is None or bool(torch.isnan(pad).all()), "leading-dimension padding changed"
    assert np.array_equal(vback, v), "v changed"


def test_update_reference_fixture():
    """Test/cholesky.cpp:146-167 restated: n = 123, v_i = 0.1/123*i - 0.5,
    alpha 0.1, beta 0.5, compared with the factor of the updated matrix to
    1e-11 percent."""
    n = 123
    A = fixtures.create_symm(n)
    v = 0.1 / n * np.arange(n) - 0.5
    L = np.linalg.cholesky(A)
    got, info, _, _ = run_update(L.copy(), 0.1, 0.5, v, False)
    R = np.linalg.cholesky(0.1 * A + 0.5 * np.outer(v, v))
    assert info == 0
    lo = np.tril_indices(n)
    assert np.allclose(got[lo], R[lo], rtol=1e-13, atol=0)


@pytest.mark.parametrize("upper", [False, True])
def test_update_large_strided(upper):
    """Several tile rows with a partial last one (n = 4097), v at stride 3."""
    n = 4097
    A, v = case(n, seed=11)
    L = np.linalg.cholesky(A)
    got, info, _, _ = run_update(stored(L, upper, np.zeros((n, n))), 1.7, 0.6, v, upper, incv=3)
    ref, rinfo = oracle_update(L, 1.7, 0.6, v)
    assert info == rinfo == 0
    Lg = got.T if upper else got
    assert np.max(np.abs(np.tril(Lg) - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("upper", [False, True])
def test_update_indefinite_leaves_factor(upper):
    """A downdate that makes the matrix indefinite: info is the oracle's
    failing column + 1 and the factor is unchanged."""
    n = 700
    A, v = case(n, seed=3)
    L = np.linalg.cholesky(A)
    F = stored(L, upper, np.full((n, n), -1.5))
    got, info, _, _ = run_update(F, 1.0, -10.0, v, upper)
    _, rinfo = oracle_update(L, 1.0, -10.0, v)
    assert rinfo > 0 and info == rinfo
    assert np.array_equal(got, F)


def test_update_beta_zero_scales_storage():
    n = 300
    A, v = case(n, seed=4)
    F = stored(np.linalg.cholesky(A), False, np.full((n, n), 2.0))
    got, info, pad, _ = run_update(F, 2.25, 0.0, v, False)
    assert info == 0
    assert np.array_equal(got, F * 1.5)
    assert bool(torch.isnan(pad).all())


def test_update_empty():
    info = torch.full((1,), -7, dtype=torch.int64, device="cuda")
    L = torch.zeros((1, 1), dtype=torch.float64, device="cuda")[:0, :0]
    la.chol_update(handle(), L, 1.0, 1.0, torch.zeros(1, dtype=torch.float64, device="cuda")[:0], info=info)
    torch.cuda.synchronize()
    assert int(info.item()) == 0


@pytest.mark.parametrize("upper", [False, True])
def test_supdate_fp32(upper):
    n = 900
    A, v = case(n, seed=9)
    L = np.linalg.cholesky(A)
    got, info, _, _ = run_update(stored(L, upper, np.zeros((n, n))), 1.3, 0.4, v, upper, dtype=torch.float32)
    ref, _ = oracle_update(L, 1.3, 0.4, v)
    assert info == 0
    Lg = got.T if upper else got
    assert np.max(np.abs(np.tril(Lg) - ref)) <= 2e-4 * np.max(np.abs(ref))


def test_update_rejects_bad_arguments():
    h = handle()
    L = torch.eye(4, dtype=torch.float64, device="cuda")
    v = torch.ones(4, dtype=torch.float64, device="cuda")
    info = torch.zeros(1, dtype=torch.int64, device="cuda")
    lib = capi.load()
    import ctypes as C
    st = lib.rmh_dchol_update(h.value, 0, 4, 1.0, 1.0, C.c_void_p(L.data_ptr()), 3, C.c_void_p(v.data_ptr()), 1,
                              C.c_void_p(info.data_ptr()), None)
    assert st == capi.INVALID_VALUE
    st = lib.rmh_dchol_update(h.value, 0, 4, 1.0, 1.0, C.c_void_p(L.data_ptr()), 4, C.c_void_p(v.data_ptr()), 0,
                              C.c_void_p(info.data_ptr()), None)
    assert st == capi.INVALID_VALUE
    st = lib.rmh_dchol_update(h.value, 0, 4, 1.0, 1.0, C.c_void_p(L.data_ptr()), 4, C.c_void_p(v.data_ptr()), 1,
                              None, None)
    assert st == capi.INVALID_VALUE
