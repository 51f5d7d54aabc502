"""Solve output, positional-compatible with the reference's ``SolveResult``.

The first seven positional/keyword parameters and their meaning follow
``solver/lp/simplex_dense.py:36-62``: ``status``, ``x`` (original variable
space), ``objective`` (original sense), ``iterations``, ``message``,
``basis``, ``gaps``.  The PDHG engine additionally fills keyword-only fields
the reference never returned: the dual vector ``y`` (rows ordered
``[A_ub; A_eq]``), the relative KKT error and its parts measured by the
reference formula (``solver/lp/pdhg.py:184-196``) on the reference-form lift,
the restart count and the primal weight at exit, and for status ``"unbounded"``
the ray of decrease that proves it (``x`` is then None and the objective
``-inf``, as the reference's simplex reports it).  Callers that only use the
reference fields see no difference.
"""

from __future__ import annotations

from typing import Optional

import numpy as np


class SolveResult:
    def __init__(
        self,
        status: str,
        x: Optional[np.ndarray],
        objective: float,
        iterations: int,
        message: str = "",
        basis: Optional[np.ndarray] = None,
        gaps: Optional[list] = None,
        *,
        y: Optional[np.ndarray] = None,
        rel_kkt: Optional[float] = None,
        rp: Optional[float] = None,
        rd: Optional[float] = None,
        gap: Optional[float] = None,
        restarts: int = 0,
        step_scale: Optional[float] = None,
        ray: Optional[np.ndarray] = None,
    ):
        self.status = status
        self.x = x
        self.objective = objective
        self.iterations = iterations
        self.message = message
        self.basis = basis
        self.gaps = gaps or []
        self.y = y
        self.rel_kkt = rel_kkt
        self.rp = rp
        self.rd = rd
        self.gap = gap
        self.restarts = restarts
        self.step_scale = step_scale    # primal weight at exit: pass it back to resume a related solve
        self.ray = ray                  # "unbounded": direction d, max |d_j| = 1, with c'd < 0

    def __repr__(self) -> str:
        obj = f"{self.objective:.6g}" if self.objective is not None else "N/A"
        return f"SolveResult(status={self.status!r}, obj={obj}, iters={self.iterations})"
