import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libpdlp_hip.so")


@pytest.fixture(scope="session")
def gpu():
    """The HIP device for GPU tests; fails (never skips) when it or the library is missing."""
    import torch
    from sovereign_solver_amd.engine._lib import lib
    assert torch.cuda.is_available(), "gpu test run without a visible HIP device"
    lib()
    return torch.device("cuda", 0)
