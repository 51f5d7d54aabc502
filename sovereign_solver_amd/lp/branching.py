"""Strong branching on the batched engine: all child LPs in one solve.

The reference scores branching candidates by solving two child LP
relaxations per candidate one after another with dual simplex
(``solver/milp/branching.py:158-203``, children built at :105-128, gains at
:131-142).  The children differ from the parent only in one bound, so they
are one shared-matrix batch: :func:`branch_strong_batched` builds the 2·K
children of the top-K most fractional candidates, solves them in one call of
:func:`~sovereign_solver_amd.lp.batch.solve_lp_pdhg_batch` (family engine),
each warm-started from the parent's primal/dual pair and primal weight, and
returns what the reference returns: ``(branch_var, down_gain, up_gain)``.

Gains follow the reference's ``_lp_gain`` (:131-142): ``inf`` for an
infeasible child (here: a Farkas certificate found by the engine), the
objective increase for an optimal one, 0 otherwise.  PDHG objectives carry
the solve tolerance (relative KKT 1e-4 by default), so gains below that
scale are noise.
"""

from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

from ..problem import Problem
from .batch import solve_lp_pdhg_batch

INTEGER_TOL = 1e-5                 # solver/config.py:14
STRONG_BRANCH_CANDIDATES = 20      # solver/config.py:44


def fractional_vars(x: np.ndarray, mask: np.ndarray) -> np.ndarray:
    """Integer-constrained variables with a fractional value (branching.py:97-101)."""
    iv = np.flatnonzero(mask)
    frac = np.abs(x[iv] - np.round(x[iv]))
    return iv[frac > INTEGER_TOL]


def node_problem(problem: Problem, lb: np.ndarray, ub: np.ndarray) -> Problem:
    """The LP relaxation of a branch-and-bound node: the problem with bounds
    ``lb``/``ub`` and no integer restrictions (branch_and_bound.py:95-113).  It
    keeps the problem's matrix objects, so the relaxations of one tree form a
    shared-matrix batch."""
    return Problem(c=problem.c, A_ub=problem.A_ub, b_ub=problem.b_ub, A_eq=problem.A_eq, b_eq=problem.b_eq,
                   lb=np.asarray(lb, float), ub=np.asarray(ub, float),
                   integer_mask=np.zeros(problem.n_vars, dtype=bool), sense=problem.sense, name=problem.name)


def child_problem(problem: Problem, lb_extra: np.ndarray, ub_extra: np.ndarray, var: int, direction: str,
                  x_val: float) -> Problem:
    """The LP relaxation of one child (branching.py:105-128)."""
    lb = np.maximum(problem.lb, lb_extra).astype(float)
    ub = np.minimum(problem.ub, ub_extra).astype(float)
    if direction == "down":
        ub[var] = min(ub[var], np.floor(x_val))
    else:
        lb[var] = max(lb[var], np.ceil(x_val))
    return node_problem(problem, lb, ub)


def _gain(r, parent_obj: float) -> float:
    if r.status == "infeasible":
        return np.inf
    if r.status == "optimal":
        return max(0.0, r.objective - parent_obj)
    return 0.0


def branch_strong_batched(
    x: np.ndarray,
    mask: np.ndarray,
    problem: Problem,
    lb_extra: np.ndarray,
    ub_extra: np.ndarray,
    parent_obj: float,
    pseudocosts=None,
    *,
    y: Optional[np.ndarray] = None,
    step_scale: Optional[float] = None,
    candidates: int = STRONG_BRANCH_CANDIDATES,
    tol: float = 1e-4,
    max_iters: int = 20_000,
) -> Tuple[int, float, float]:
    """Strong branching with every child LP in one batched solve.

    Same arguments and result as the reference's ``branch_strong``
    (``branching.py:158-171``); ``y`` and ``step_scale`` (the parent's dual
    and primal weight, e.g. from its :class:`SolveResult`) warm-start the
    children.  ``pseudocosts``, when given, receives ``update_down`` /
    ``update_up`` calls as in the reference.
    """
    cands = fractional_vars(np.asarray(x, float), np.asarray(mask, bool))
    if cands.size == 0:
        return -1, 0.0, 0.0
    frac = np.abs(x[cands] - np.round(x[cands]))
    score = np.minimum(frac, 1.0 - frac)
    top = cands[np.argsort(score)[::-1][: min(candidates, cands.size)]]
    children = []
    for j in top:
        children.append(child_problem(problem, lb_extra, ub_extra, int(j), "down", float(x[j])))
        children.append(child_problem(problem, lb_extra, ub_extra, int(j), "up", float(x[j])))
    warm = [(np.clip(x, ch.lb, ch.ub), y, step_scale) for ch in children]
    res = solve_lp_pdhg_batch(children, warm_start=warm, tol=tol, max_iters=max_iters, engine="family")
    best = (int(top[0]), 0.0, 0.0)
    best_score = -np.inf
    for k, j in enumerate(top):
        down, up = _gain(res[2 * k], parent_obj), _gain(res[2 * k + 1], parent_obj)
        s = max(1e-6, down) * max(1e-6, up)               # branching.py:196-198
        if pseudocosts is not None:
            pseudocosts.update_down(int(j), down)
            pseudocosts.update_up(int(j), up)
        if s > best_score:
            best_score, best = s, (int(j), down, up)
    return best
