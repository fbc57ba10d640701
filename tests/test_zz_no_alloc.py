"""rmh_dpotrf_noalloc: rmh_dpotrf on the library's static per-device scratch,
allocating nothing.  Checked through the C-ABI against rmh_dpotrf on the same
inputs, bit for bit (both run the same kernels; the no-allocation path runs
on one stream, and the single-stream schedule is bit-identical to the
look-ahead one): two calls in a row, the second smaller and in the other
storage triangle, with and without a caller info word; free device memory is
unchanged across them on a handle that has never allocated a workspace; and
orders beyond the scratch are rejected before anything runs."""
import ctypes as C

import numpy as np
import pytest
import torch

from remora_amd import capi, synth
from remora_amd import linalg as la
from tests.gpu_util import dev, host

pytestmark = pytest.mark.gpu


def _noalloc(h, A, upper=False, info=None):
    n = A.shape[0]
    return capi.load().rmh_dpotrf_noalloc(h.value, int(upper), n, C.c_void_p(A.data_ptr()), max(1, A.stride(1)),
                                          None if info is None else C.c_void_p(info.data_ptr()),
                                          C.c_void_p(torch.cuda.current_stream().cuda_stream))


def test_noalloc_matches_potrf_twice_and_allocates_nothing():
    print("noalloc: n=2000 lower, then n=700 upper", flush=True)
    h_ref = capi.Handle(0)
    h_na = capi.Handle(0)  # fresh: no workspace allocated yet
    try:
        cases = [(2000, False, 7), (700, True, 8)]
        refs, inputs = [], []
        for n, upper, seed in cases:
            A = synth.gram_spd(n, seed=seed)
            inputs.append(A)
            dR = dev(A, ld=n + 3)
            assert int(la.potrf(h_ref, dR, upper=upper).item()) == 0
            refs.append(host(dR))
        # every tensor of the no-allocation calls exists before the measurement
        dAs = [dev(A, ld=A.shape[0] + 3) for A in inputs]
        infos = [torch.full((1,), 99, dtype=torch.int64, device="cuda") for _ in cases]
        dA_noinfo = dev(inputs[1], ld=inputs[1].shape[0] + 3)
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info()
        for (n, upper, _), dA, info in zip(cases, dAs, infos):
            capi.check(_noalloc(h_na, dA, upper, info))
        capi.check(_noalloc(h_na, dA_noinfo, True, None))  # info in the scratch
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info()
        assert free0 - free1 < (1 << 20), (free0, free1)
        for (n, upper, _), dA, info, ref in zip(cases, dAs, infos, refs):
            assert int(info.item()) == 0
            assert np.array_equal(host(dA), ref), n
        assert np.array_equal(host(dA_noinfo), refs[1])
    finally:
        h_na.close()
        h_ref.close()


def test_noalloc_reports_info_like_potrf():
    """A leading minor that is not positive definite: the same LAPACK info."""
    n, k = 900, 333
    A = synth.gram_spd(n, seed=1)
    L = np.linalg.cholesky(A[:k + 1, :k + 1])
    A[k, k] -= L[k, k] ** 2 * 1.5
    h = capi.Handle(0)
    try:
        info = torch.zeros(1, dtype=torch.int64, device="cuda")
        capi.check(_noalloc(h, dev(A), False, info))
        assert int(info.item()) == k + 1
    finally:
        h.close()


def test_noalloc_rejects_orders_beyond_the_scratch():
    """n above 32768 at nb_inner 32 (16384 at 64) is an error, returned before
    any launch or memory access (the pointer below is a one-element buffer)."""
    h = capi.Handle(0)
    try:
        tiny = torch.zeros((1, 1), dtype=torch.float64, device="cuda")
        lib = capi.load()
        s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        for n, nbi in ((32769, 0), (16385, 64)):
            la.set_potrf_blocking(h, 0, nbi)
            st = lib.rmh_dpotrf_noalloc(h.value, 0, n, C.c_void_p(tiny.data_ptr()), n, None, s)
            assert st == capi.INVALID_VALUE, st
            assert b"maximum" in lib.rmh_last_error()
        la.set_potrf_blocking(h, 0, 0)
        # with the default blocking restored a small order runs
        A = synth.gram_spd(64, seed=2)
        dA = dev(A)
        assert _noalloc(h, dA) == capi.SUCCESS
        torch.cuda.synchronize()
    finally:
        h.close()
