"""Strong branching on the batched engine (needs an MI355X): the child LPs of
the reference's branch_strong (solver/milp/branching.py:158-210), solved in
one batch, must give the gains an exact LP solver (HiGHS, FP64) gives."""

import numpy as np
import pytest
from scipy.optimize import linprog

from sovereign_solver_amd import branch_strong_batched, solve_lp_pdhg
from sovereign_solver_amd.lp.branching import child_problem, fractional_vars
from sovereign_solver_amd.workloads import SetCoverNodes

pytestmark = pytest.mark.gpu


def _exact(p):
    r = linprog(p.c, A_ub=p.A_ub, b_ub=p.b_ub, bounds=list(zip(p.lb, p.ub)), method="highs-ds")
    return np.inf if r.status == 2 else r.fun


class _Pseudocosts:
    def __init__(self):
        self.down, self.up = {}, {}

    def update_down(self, j, g):
        self.down[j] = g

    def update_up(self, j, g):
        self.up[j] = g


def test_strong_branching_batched_gains(gpu):
    sc = SetCoverNodes(rows=60, cols=400, density=0.03)
    p = sc.root()
    mask = np.ones(p.n_vars, bool)
    root = solve_lp_pdhg(p, tol=1e-4, precision="fp32")
    assert root.status == "optimal"
    pc = _Pseudocosts()
    var, down, up = branch_strong_batched(root.x, mask, p, p.lb, p.ub, root.objective, pc, y=root.y,
                                          step_scale=root.step_scale, candidates=8)
    cands = fractional_vars(root.x, mask)
    assert var in set(cands.tolist()) and len(pc.down) == len(pc.up) == min(8, cands.size)
    parent = _exact(p)
    tol = 3e-3 * (1 + abs(parent))
    exact_scores = {}
    for j in pc.down:
        gd = _exact(child_problem(p, p.lb, p.ub, j, "down", root.x[j])) - parent
        gu = _exact(child_problem(p, p.lb, p.ub, j, "up", root.x[j])) - parent
        for got, want in ((pc.down[j], gd), (pc.up[j], gu)):
            if np.isinf(want):
                assert np.isinf(got)
            else:
                assert abs(got - max(0.0, want)) <= tol, (j, got, want)
        exact_scores[j] = max(1e-6, max(0.0, gd)) * max(1e-6, max(0.0, gu))
    # the chosen variable scores within the gains' tolerance of the exact best
    best = max(exact_scores.values())
    assert exact_scores[var] >= best - 2 * tol * (1 + best ** 0.5) - 1e-9


def test_strong_branching_flags_infeasible_children(gpu):
    # a row covered by a single fractional variable: its down child is infeasible
    sc = SetCoverNodes(rows=60, cols=400, density=0.03)
    p = sc.root()
    root = solve_lp_pdhg(p, tol=1e-4, precision="fp32")
    j = int(fractional_vars(root.x, np.ones(p.n_vars, bool))[0])
    row = sc.A[:, j].nonzero()[0][0]
    others = np.setdiff1d(sc.A[row].indices, [j])
    lb, ub = p.lb.copy(), p.ub.copy()
    ub[others] = 0.0                                    # only j can cover that row now
    q = sc.node_from(lb, ub)
    rq = solve_lp_pdhg(q, tol=1e-4, precision="fp32")
    mask = np.zeros(q.n_vars, bool)
    mask[j] = True
    x = rq.x.copy()
    x[j] = 0.5                                          # branch on j
    pc = _Pseudocosts()
    branch_strong_batched(x, mask, q, q.lb, q.ub, rq.objective, pc, y=rq.y, step_scale=rq.step_scale)
    assert np.isinf(pc.down[j]) and np.isfinite(pc.up[j])
