import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def handle():
    """One library handle on device 0 for the whole GPU session."""
    import torch  # noqa: F401  (shares its HIP runtime with the library)
    from remora_amd import capi
    h = capi.Handle(0)
    yield h
    h.close()
