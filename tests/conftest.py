import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls through the C-ABI")


@pytest.fixture(scope="session")
def gpu():
    """The library context on cuda:0; fails (never skips) when no device is visible."""
    import torch

    assert torch.cuda.is_available(), "GPU test collected on a host without a visible HIP device"
    import lightne_amd as L

    return L.context(0)
