"""Drop-in for the reference's GPU entry points (``solver/lp/pdhg_gpu.py``).

``get_available_device()`` (``pdhg_gpu.py:42-67``) reports ``"cuda"`` when a
HIP device is visible to PyTorch-ROCm (HIP devices surface as ``cuda`` in
torch) and ``libpdlp_hip.so`` loads, else ``"cpu"``.

``solve_lp_pdhg_gpu(problem, device="auto", max_iters, tol)``
(``pdhg_gpu.py:315-443``) runs the HIP engine.  Device strings are resolved
as in ``DeviceEngine.__init__`` (``pdhg_gpu.py:78-95``): an unknown string
raises ``ValueError``.  Unlike the reference, there is no NumPy fallback:
``"cpu"``, ``"mps"``, or ``"auto"``/``"cuda"`` without a HIP device raise
``RuntimeError`` instead of silently running elsewhere — this engine's
results always come from the GPU kernels.
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
    if device in ("cpu", "mps"):
        raise RuntimeError(f"device={device!r}: this engine runs only on a HIP GPU (no CPU/MPS path)")
    if get_available_device() != "cuda":
        raise RuntimeError("no usable HIP device (torch.cuda unavailable or libpdlp_hip.so not built)")
    return solve_lp_pdhg(problem, max_iters=max_iters, tol=tol, **kwargs)
