"""Batched drop-in: ``solve_lp_pdhg_batch(problems) -> list[SolveResult]``.

The reference has no batch call; its callers loop over ``solve_lp_pdhg``
(``solver/lp/pdhg.py:152``) or, in branch-and-bound, over simplex node solves
(``solver/milp/branch_and_bound.py:367-374``).  This call solves all LPs in one
persistent HIP launch and returns one :class:`SolveResult` per problem with the
same fields and semantics as :func:`solve_lp_pdhg` (box form, FP32 iterates,
FP64 KKT, check cadence 50; ``scheme="adaptive"`` (default) or
``"reference"`` restarts as in :func:`solve_lp_pdhg`).

``warm_start`` is an optional list (one entry per problem, ``None`` allowed)
of ``(x0, y0)`` or ``(x0, y0, step_scale)``: x0 in the original variable
space, y0 the duals of ``[A_ub; A_eq]`` (y ≤ 0 on the ``<=`` rows), and a
primal weight (τ ← ω τ, σ ← σ/ω).  A bad warm start only costs iterations.
"""

from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from ..config import MAX_PDHG_ITERS
from ..engine.batch import BatchEngine, PackedBatch
from ..problem import Problem
from ..result import SolveResult


def pack_warm_start(pk: PackedBatch, warm_start):
    if warm_start is None:
        return None, None, None
    x0 = np.zeros(pk.N)
    y0 = np.zeros(pk.M)
    w = np.full(len(pk.problems), np.nan)     # NaN: the engine's default starting weight
    for k, ws in enumerate(warm_start):
        if ws is None:
            continue
        p = pk.problems[k]
        xs, ys = ws[0], ws[1]
        if xs is not None:
            x0[pk.col_off[k]: pk.col_off[k] + pk.ns[k]] = np.clip(np.asarray(xs, float), p.lb, p.ub)
        if ys is not None:
            yk = np.asarray(ys, float).copy()
            yk[: p.n_ineq] = np.minimum(yk[: p.n_ineq], 0.0)
            y0[pk.row_off[k]: pk.row_off[k] + pk.ms[k]] = yk
        if len(ws) > 2 and ws[2] is not None:
            w[k] = float(ws[2])
    return x0, y0, w


def results_of(engine: BatchEngine, out, xs, ys) -> list:
    res = []
    for k, (status, x, y, rec) in enumerate(engine.unpack(out, xs, ys)):
        p = engine.pk.problems[k]
        x = np.clip(x, np.where(np.isfinite(p.lb), p.lb, -1e30), np.where(np.isfinite(p.ub), p.ub, 1e30))
        obj = float(np.asarray(p.c, float) @ x)
        if p.sense == "max":
            obj = -obj
        res.append(SolveResult(status, x.copy(), obj, int(rec["iterations"]), y=y.copy(),
                               rel_kkt=float(rec["rel_kkt"]), rp=float(rec["rp"]), rd=float(rec["rd"]),
                               gap=float(rec["gap"]), restarts=int(rec["restarts"]),
                               step_scale=float(rec["weight"])))
    return res


def solve_lp_pdhg_batch(
    problems: Sequence[Problem],
    warm_start: Optional[Sequence] = None,
    tol: float = 1e-4,
    max_iters: int = MAX_PDHG_ITERS,
    *,
    check_every: int = 50,
    block: int = 0,
    scheme: str = "adaptive",
) -> list:
    problems = list(problems)
    if not problems:
        return []
    trivial = [k for k, p in enumerate(problems) if p.n_vars == 0 or p.n_constraints == 0]
    live = [k for k in range(len(problems)) if k not in set(trivial)]
    out = [None] * len(problems)
    for k in trivial:                                    # pdhg.py:171-172
        out[k] = SolveResult("optimal", np.zeros(problems[k].n_vars), 0.0, 0)
    if live:
        pk = PackedBatch([problems[k] for k in live])
        eng = BatchEngine(pk, block=block, scheme=scheme)
        ws = None if warm_start is None else [warm_start[k] for k in live]
        x0, y0, w = pack_warm_start(pk, ws)
        o, xs, ys = eng.solve(tol, max_iters, check_every, x0, y0, w)
        for k, r in zip(live, results_of(eng, o, xs, ys)):
            out[k] = r
    return out
