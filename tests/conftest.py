import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X GPU (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library on a GPU; fails loudly when either is missing."""
    import torch
    from advection_miniapp_amd import _lib
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    return _lib.load()
