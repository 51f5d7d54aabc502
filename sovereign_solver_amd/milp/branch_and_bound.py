"""Branch-and-bound whose node LPs are solved in batches on the GPU.

The reference's branch-and-cut (``solver/milp/branch_and_bound.py:156-470``)
takes one node at a time from a best-bound heap and solves its LP relaxation
with dual simplex (:363-374).  Node LPs differ from their parent only in
variable bounds, so here each round takes up to ``batch`` best-bound nodes,
solves all their relaxations in one call of the shared-matrix batched engine
(:func:`~sovereign_solver_amd.lp.batch.solve_lp_pdhg_batch`, family engine;
each node warm-started from its parent's primal/dual pair and primal weight)
and then processes the results as the reference does one by one: prune
infeasible nodes (Farkas certificates from the engine) and nodes whose bound
cannot beat the incumbent, accept integer-feasible solutions, branch on the
most fractional variable (:421-441) otherwise.

What is not here: the reference's presolve, cutting planes, primal
heuristics and pseudocost/strong branching inside the loop (strong branching
in batches is :func:`~sovereign_solver_amd.lp.branching.branch_strong_batched`).

PDHG solves each node LP to a relative KKT tolerance (``lp_tol``), so node
bounds are approximate: a node is pruned when its LP objective, less a margin
of ``2 lp_tol (1 + |obj|)`` for the objective's own error, is within
``rel_gap`` (relative, default 1e-4) of the incumbent, and the result is
optimal to that gap.  Integer-feasible candidates are rounded and checked
against the original rows and bounds before they become the incumbent, so
the returned x is always feasible to ``FEASIBILITY_TOL``-scaled checks.

A node whose LP ends unsolved (iteration limit or a numerical error) is never
dropped silently: its rounded point is still offered as an incumbent, a
finite fractional point is branched on, and otherwise (an integral or
non-finite point says nothing about the subtree) the node is solved again from
a cold start with four times the iteration limit, twice at most.  A node still
unsolved after that leaves its subtree unexplored, and the search then ends
``node_limit`` rather than ``optimal`` or ``infeasible``.
"""

from __future__ import annotations

import dataclasses
import heapq
import time
from typing import List, Optional

import numpy as np

from ..lp.batch import solve_lp_pdhg_batch
from ..lp.branching import node_problem
from ..problem import Problem

# solver/config.py:10,14,42-43
FEASIBILITY_TOL = 1e-6
INTEGER_TOL = 1e-5
MAX_BB_NODES = 1_000_000
BB_TIME_LIMIT_S = 300.0


@dataclasses.dataclass
class MILPSolveResult:
    """Result of a MILP solve (branch_and_bound.py:82-90)."""
    status: str                 # 'optimal' | 'infeasible' | 'node_limit' | 'time_limit'
    x: Optional[np.ndarray]
    objective: float
    nodes: int                  # node LPs solved
    lp_relaxation: float = np.inf
    gap: float = 0.0


@dataclasses.dataclass
class _Node:
    lb: np.ndarray
    ub: np.ndarray
    bound: float                # parent's LP objective (min form)
    depth: int
    warm: Optional[tuple]       # parent's (x, y, primal weight)
    tries: int = 0              # earlier solves of this node that ended unsolved


def _min_obj(p: Problem, x: np.ndarray) -> float:
    return float(np.dot(p.c, x))


def _feasible(p: Problem, x: np.ndarray, tol: float) -> bool:
    """Rows and bounds of ``p`` at x within tol (scaled by the row's size)."""
    if p.n_ineq:
        r = p.A_ub @ x - p.b_ub
        if np.any(r > tol * (1.0 + np.abs(p.b_ub))):
            return False
    if p.n_eq:
        r = np.abs(p.A_eq @ x - p.b_eq)
        if np.any(r > tol * (1.0 + np.abs(p.b_eq))):
            return False
    return bool(np.all(x >= p.lb - tol) and np.all(x <= p.ub + tol))


def _rounded(p: Problem, x: np.ndarray) -> np.ndarray:
    xr = np.clip(x, p.lb, p.ub)
    m = p.integer_mask
    xr[m] = np.round(xr[m])
    return xr


def _gap(incumbent: float, bound: float) -> float:
    """branch_and_bound.py:147-151."""
    if not np.isfinite(incumbent):
        return np.inf
    return abs(incumbent - bound) / (1.0 + abs(incumbent))


def solve_milp_batched(
    problem: Problem,
    time_limit: float = BB_TIME_LIMIT_S,
    node_limit: int = MAX_BB_NODES,
    *,
    batch: int = 256,
    lp_tol: float = 1e-4,
    rel_gap: float = 1e-4,
    int_tol: float = 1e-3,
    max_lp_iters: int = 20_000,
    verbose: bool = False,
) -> MILPSolveResult:
    """Best-bound branch-and-bound with batched node relaxations.

    Same call shape and result as the reference's ``solve_milp``
    (``branch_and_bound.py:156-165``) for the search itself: ``time_limit``,
    ``node_limit``, ``verbose``; ``batch`` nodes are solved per round.
    ``int_tol`` is the distance to an integer below which a PDHG value
    (accurate to about ``lp_tol``) counts as integral; such candidates are
    rounded and verified before they are accepted.
    """
    t0 = time.monotonic()
    p = problem
    mask = np.asarray(p.integer_mask, bool)
    sign = -1.0 if p.sense == "max" else 1.0

    def result(status, x, obj_min, nodes, root_bound, best_bound):
        obj = sign * obj_min if np.isfinite(obj_min) else (np.inf if status == "infeasible" else obj_min)
        return MILPSolveResult(status, x, obj, nodes, sign * root_bound if np.isfinite(root_bound) else root_bound,
                               _gap(obj_min, best_bound))

    lb0 = np.asarray(p.lb, float).copy()
    ub0 = np.asarray(p.ub, float).copy()
    # the root, through the same engine (its Farkas check reports an infeasible relaxation)
    root = solve_lp_pdhg_batch([node_problem(p, lb0, ub0)], tol=lp_tol, max_iters=max_lp_iters,
                               engine="family")[0]
    nodes = 1
    if root.status == "infeasible":
        return result("infeasible", None, np.inf, nodes, np.inf, np.inf)
    root_bound = _min_obj(p, root.x) if root.x is not None and np.all(np.isfinite(root.x)) else -np.inf
    incumbent_x, incumbent = None, np.inf

    def consider(x: np.ndarray):
        nonlocal incumbent_x, incumbent
        xr = _rounded(p, x)
        if _feasible(p, xr, 10 * FEASIBILITY_TOL + lp_tol):
            v = _min_obj(p, xr)
            if v < incumbent - 1e-9 * (1 + abs(v)):
                incumbent, incumbent_x = v, xr
                if verbose:
                    print(f"  [incumbent] nodes {nodes} obj {sign * v:.6f}")

    def cutoff() -> float:
        """Bounds at or above this cannot improve the incumbent by more than rel_gap."""
        return incumbent - rel_gap * (1 + abs(incumbent)) if np.isfinite(incumbent) else np.inf

    heap: List = []
    counter = 0

    def push(node: _Node):
        nonlocal counter
        counter += 1
        heapq.heappush(heap, (node.bound, counter, node))

    ivars = np.flatnonzero(mask)

    def branch(node_lb, node_ub, x, y, w, obj, depth):
        """Most fractional integer variable (branch_and_bound.py:421-441)."""
        frac = np.abs(x[ivars] - np.round(x[ivars]))
        if frac.size == 0 or frac.max() <= int_tol:
            return
        j = int(ivars[np.argmax(frac)])                 # |x - round(x)| is largest at 0.5
        lo_ub = node_ub.copy()
        lo_ub[j] = min(lo_ub[j], np.floor(x[j]))
        hi_lb = node_lb.copy()
        hi_lb[j] = max(hi_lb[j], np.ceil(x[j]))
        for lbc, ubc in ((node_lb, lo_ub), (hi_lb, node_ub)):
            if np.all(lbc <= ubc + FEASIBILITY_TOL):
                push(_Node(lbc, ubc, obj, depth + 1, (x, y, w)))

    unresolved = 0                                      # subtrees left unexplored (unsolved node LPs)

    def settle(node_lb, node_ub, r, depth, parent_bound, tries=0):
        """Process one solved node relaxation."""
        nonlocal unresolved
        if r.status == "infeasible":
            return
        x = r.x
        finite = x is not None and bool(np.all(np.isfinite(x)))
        if r.status == "optimal" and finite:
            obj = _min_obj(p, x)
            consider(x)
            if obj - 2.0 * lp_tol * (1.0 + abs(obj)) >= cutoff():
                return                                  # cannot beat the incumbent
            frac = np.abs(x[ivars] - np.round(x[ivars]))
            if frac.size == 0 or frac.max() <= int_tol:
                return                                  # integral: handled by consider()
            branch(node_lb, node_ub, x, r.y, r.step_scale, obj, depth)
            return
        # not solved: keep the parent's bound; a finite fractional point still splits the node
        if finite:
            consider(x)
            frac = np.abs(x[ivars] - np.round(x[ivars]))
            if frac.size and frac.max() > int_tol:
                branch(node_lb, node_ub, x, r.y, r.step_scale, parent_bound, depth)
                return
        if tries < 2:                                   # solve it again, cold, with more iterations
            push(_Node(node_lb, node_ub, parent_bound, depth, None, tries + 1))
        else:
            unresolved += 1

    # an unsolved root bounds nothing: its subtree keeps the trivial bound
    settle(lb0, ub0, root, 0, root_bound if root.status == "optimal" else -np.inf)
    while heap and nodes < node_limit:
        if time.monotonic() - t0 > time_limit:
            best = heap[0][0] if heap else incumbent
            return result("time_limit", incumbent_x, incumbent, nodes, root_bound, min(best, incumbent))
        take: List[_Node] = []
        while heap and len(take) < min(batch, node_limit - nodes):
            bound, _, node = heapq.heappop(heap)
            if bound < cutoff():
                take.append(node)
        if not take:
            break
        # nodes solved again after an unsolved attempt get four times the iterations per attempt
        for tries in sorted({nd.tries for nd in take}):
            group = [nd for nd in take if nd.tries == tries]
            children = [node_problem(p, nd.lb, nd.ub) for nd in group]
            warm = [(np.clip(nd.warm[0], c.lb, c.ub), nd.warm[1], nd.warm[2]) if nd.warm else None
                    for nd, c in zip(group, children)]
            res = solve_lp_pdhg_batch(children, warm_start=warm if any(warm) else None, tol=lp_tol,
                                      max_iters=max_lp_iters * 4 ** tries, engine="family")
            nodes += len(group)
            for nd, r in zip(group, res):
                settle(nd.lb, nd.ub, r, nd.depth, nd.bound, nd.tries)
        if verbose:
            best = heap[0][0] if heap else incumbent
            print(f"  nodes {nodes} open {len(heap)} incumbent {sign * incumbent:.6f} bound {sign * best:.6f}")

    live = [e[0] for e in heap if e[0] < cutoff()]       # open nodes that could still improve
    if live:
        return result("node_limit", incumbent_x, incumbent, nodes, root_bound, min(min(live), incumbent))
    if unresolved:                                       # a subtree was left unexplored: no proof
        return result("node_limit", incumbent_x, incumbent, nodes, root_bound, -np.inf)
    if incumbent_x is None:
        return result("infeasible", None, np.inf, nodes, root_bound, np.inf)
    return result("optimal", incumbent_x, incumbent, nodes, root_bound, incumbent)
