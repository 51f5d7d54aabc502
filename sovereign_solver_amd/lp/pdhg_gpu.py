"""Drop-in for the reference's GPU entry points (``solver/lp/pdhg_gpu.py``).

``get_available_device()`` (``pdhg_gpu.py:42-67``) reports ``"cuda"`` when a
HIP device is visible to PyTorch-ROCm (HIP devices surface as ``cuda`` in
torch) and ``libpdlp_hip.so`` loads, else ``"cpu"``.

``solve_lp_pdhg_gpu(problem, device="auto", max_iters, tol)``
(``pdhg_gpu.py:315-443``) runs the HIP engine.  Device strings are resolved
as in ``DeviceEngine.__init__`` (``pdhg_gpu.py:78-95``): an unknown string
raises ``ValueError``, and a requested device that this engine does not have
logs a warning and runs on the one it has — the reference falls back from an
unavailable device to the detected one the same way (``:83-89``).  This engine
has one device, the HIP GPU: ``"cpu"`` and ``"mps"`` warn and solve there (so
a caller's ``device="cpu"`` still gets a result, as the reference's
``tests/test_pdhg_gpu.py:148-157`` expects), and without a HIP device every
request raises ``RuntimeError`` — there is no NumPy path, so every result
comes from the GPU kernels.
"""

from __future__ import annotations

import logging

from ..config import MAX_PDHG_ITERS, PDHG_TOL
from ..problem import Problem
from ..result import SolveResult
from .pdhg import solve_lp_pdhg

logger = logging.getLogger(__name__)


def get_available_device() -> str:
    try:
        import torch
        if not torch.cuda.is_available():
            return "cpu"
        from ..engine._lib import lib
        lib()
        return "cuda"
    except Exception:                                   # no torch, no device or no library
        return "cpu"


def solve_lp_pdhg_gpu(
    problem: Problem,
    device: str = "auto",
    max_iters: int = MAX_PDHG_ITERS,
    tol: float = PDHG_TOL,
    **kwargs,
) -> SolveResult:
    if device not in ("auto", "cuda", "mps", "cpu"):
        raise ValueError(f"Unknown device: {device!r}. Choose 'auto', 'cuda', 'mps', or 'cpu'.")
    if get_available_device() != "cuda":
        raise RuntimeError("no usable HIP device (torch.cuda unavailable or libpdlp_hip.so not built); "
                           "this engine has no CPU path")
    if device in ("cpu", "mps"):
        logger.warning(f"Requested device {device!r}: this engine runs on the HIP GPU only; "
                       f"solving on 'cuda' instead.")
    return solve_lp_pdhg(problem, max_iters=max_iters, tol=tol, **kwargs)
