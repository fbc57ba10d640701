"""Runs the C++ front-end test program (tests/cpp/test_frontend.cpp), which
restates the reference's HIP tests on the header-only hip_tag front-end."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_frontend")


def test_cpp_frontend():
    assert os.path.exists(BIN), "tests/cpp/test_frontend not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
