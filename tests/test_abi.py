"""The C-ABI library loads and exports every symbol include/remora_hip/*.h
declares (no compute call: no GPU needed)."""
import ctypes
import glob
import os
import re

import pytest

from remora_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "remora_hip", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(rmh_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    declared = _declared()
    assert len(declared) >= 20
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    declared = _declared()
    assert declared == set(capi.SIGNATURES), declared ^ set(capi.SIGNATURES)


def test_status_strings_and_errors_without_gpu():
    lib = capi.load()
    assert lib.rmh_status_string(0) == b"success"
    assert lib.rmh_status_string(5) == b"matrix is not positive definite"
    h = ctypes.c_void_p()
    st = lib.rmh_create(ctypes.byref(h), 0)
    if st == capi.SUCCESS:  # a GPU is present: nothing more to check here
        lib.rmh_destroy(h)
        pytest.skip("GPU present")
    assert st in (capi.HIP_ERROR, capi.INVALID_VALUE)
    assert len(lib.rmh_last_error()) > 0
    assert lib.rmh_dgemm(None, 0, 0, 1, 1, 1, 1.0, None, 1, None, 1, 0.0, None, 1, None) == capi.INVALID_HANDLE
