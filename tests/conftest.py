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


def pytest_terminal_summary(terminalreporter):
    """Objective checks that needed more than the 1e-4 bar (tests/parity.py)."""
    try:
        from tests.parity import FALLBACKS
    except Exception:                                    # pragma: no cover
        return
    if FALLBACKS:
        terminalreporter.write_line(f"objective parity: {len(FALLBACKS)} results beyond 1e-4 of HiGHS:")
        for rec in FALLBACKS:
            terminalreporter.write_line("  " + " ".join(f"{v:.3g}" if isinstance(v, float) else str(v) for v in rec))
