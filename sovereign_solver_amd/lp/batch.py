"""Batched drop-in: ``solve_lp_pdhg_batch(problems) -> list[SolveResult]``.

The reference has no batch call; its callers loop over ``solve_lp_pdhg``
(``solver/lp/pdhg.py:152``) or, in branch-and-bound, over simplex node solves
(``solver/milp/branch_and_bound.py:367-374``).  This call solves all LPs in one
batched solve and returns one :class:`SolveResult` per problem with the same
fields and semantics as :func:`solve_lp_pdhg` (box form, FP32 iterates — FP64
for the small LPs of the persistent engine —, FP64 KKT, check cadence 50;
``scheme="adaptive"`` (default) or ``"reference"`` restarts as in
:func:`solve_lp_pdhg`).

``engine``: ``"persistent"`` (one workgroup per LP, any mix of matrices; an
LP needs 12n+8m+8 bytes of LDS), ``"family"`` (all LPs share one constraint
matrix and step in lockstep, 64 LPs per wavefront; no size limit) or
``"auto"`` (family when every problem holds the same ``A_ub``/``A_eq``
objects and there are at least 16 of them, or when an LP is too large for
LDS; else persistent).

``scaling`` (persistent engine): diagonal preconditioning of every LP
(``"ruiz"``, ``"geometric"``, ``"pock_chambolle"``, ``"ruiz_pc"``; see
:func:`~sovereign_solver_amd.lp.pdhg.solve_lp_pdhg`); the tolerance applies to
the original LPs' KKT and the results come back unscaled.

``tune=True``: the launch configuration comes from the autotuner
(:mod:`~sovereign_solver_amd.engine.tune`): for the persistent engine the
check cadence, each part's workgroup size, the FP32 part's lines per round and
the FP32/FP64 split (:func:`~sovereign_solver_amd.engine.tune.tune_batch`);
for the family engine the launch shape and the check cadence
(:func:`~sovereign_solver_amd.engine.tune.tune_family`).  A shape bucket seen
before takes its cached choice; a new one is tuned by solving this batch once
per candidate (from the given warm start).

``warm_start`` is an optional list (one entry per problem, ``None`` allowed)
of ``(x0, y0)`` or ``(x0, y0, step_scale)``: x0 in the original variable
space, y0 the duals of ``[A_ub; A_eq]`` (y ≤ 0 on the ``<=`` rows), and a
primal weight (τ ← ω τ, σ ← σ/ω).  A bad warm start only costs iterations.
"""

from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from ..config import MAX_PDHG_ITERS
from ..engine import tune as T
from ..engine.batch import BatchEngine, PackedBatch, lds_bytes_of
from ..engine.family import FamilyBatch, FamilyEngine
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


def family_warm_start(fb: FamilyBatch, warm_start):
    """Per-LP warm starts as [B, n] / [B, m] arrays and weights (NaN = default)."""
    if warm_start is None:
        return None, None, None
    X = np.stack([np.clip(0.0, p.lb, p.ub) for p in fb.problems])     # the problems' own column order
    Y = np.zeros((fb.count, fb.m))
    w = np.full(fb.count, np.nan)
    for k, ws in enumerate(warm_start):
        if ws is None:
            continue
        if ws[0] is not None:
            X[k] = np.asarray(ws[0], float)
        if ws[1] is not None:
            Y[k] = np.asarray(ws[1], float)
        if len(ws) > 2 and ws[2] is not None:
            w[k] = float(ws[2])
    return X, Y, w


def results_of(engine, out, xs, ys) -> list:
    """SolveResults of a persistent (:class:`BatchEngine`) or family engine run."""
    problems = engine.pk.problems if isinstance(engine, BatchEngine) else engine.fb.problems
    res = []
    for k, (status, x, y, rec) in enumerate(engine.unpack(out, xs, ys)):
        p = problems[k]
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
    engine: str = "auto",
    scaling: str = "none",
    tune: bool = False,
) -> list:
    if engine not in ("auto", "persistent", "family"):
        raise ValueError("engine must be 'auto', 'persistent' or 'family'")
    if scaling != "none" and engine == "family":
        raise ValueError("scaling is supported by the persistent engine (engine='persistent' or 'auto')")
    problems = list(problems)
    if not problems:
        return []
    trivial = [k for k, p in enumerate(problems) if p.n_vars == 0 or p.n_constraints == 0]
    live = [k for k in range(len(problems)) if k not in set(trivial)]
    out = [None] * len(problems)
    for k in trivial:                                    # pdhg.py:171-172
        out[k] = SolveResult("optimal", np.zeros(problems[k].n_vars), 0.0, 0)
    if live and engine == "auto":
        ps = [problems[k] for k in live]
        same = all(p.A_ub is ps[0].A_ub and p.A_eq is ps[0].A_eq for p in ps)
        too_big = max(lds_bytes_of(p.n_constraints, p.n_vars) for p in ps) > 160 * 1024
        engine = "family" if scaling == "none" and ((same and len(ps) >= 16) or (too_big and same)) else "persistent"
    if live and engine == "family":
        fb = FamilyBatch([problems[k] for k in live])
        eng = FamilyEngine(fb, scheme=scheme, check_every=check_every)
        if fb.count >= 256 or tune:              # launch shape by timing; a few ms per engine
            eng.autotune(tol)
        ws = None if warm_start is None else [warm_start[k] for k in live]
        X, Y, w = family_warm_start(fb, ws)
        if tune:
            T.tune_family(eng, tol, max_iters, X, Y, w)
        o, xs, ys = eng.solve(tol, max_iters, X, Y, w)
        for k, r in zip(live, results_of(eng, o, xs, ys)):
            out[k] = r
    elif live:
        pk = PackedBatch([problems[k] for k in live], scaling=scaling)
        eng = BatchEngine(pk, block=block, scheme=scheme, check_every=check_every)
        ws = None if warm_start is None else [warm_start[k] for k in live]
        x0, y0, w = pack_warm_start(pk, ws)
        if tune:
            T.tune_batch(eng, tol, max_iters, x0, y0, w)
        o, xs, ys = eng.solve(tol, max_iters, None, x0, y0, w)
        for k, r in zip(live, results_of(eng, o, xs, ys)):
            out[k] = r
    return out
