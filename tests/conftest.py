import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm device) and the built HIP library")


@pytest.fixture
def hip_backend():
    """Activate the hip backend on the GPU for one test."""
    from pytme_amd.backends import backend as be

    be.change_backend("hip", device="cuda")
    yield be
    be.change_backend("hip", device="cuda")


@pytest.fixture
def cpu_backend():
    """The hip backend on the CPU device (host logic only)."""
    from pytme_amd.backends import backend as be

    be.change_backend("hip", device="cpu")
    yield be
    be.change_backend("hip", device="cuda")
