"""Host-side pieces of the batched branch-and-bound (no GPU)."""

import numpy as np
import scipy.sparse as sp

from sovereign_solver_amd.lp.branching import child_problem, node_problem
from sovereign_solver_amd.milp.branch_and_bound import _feasible, _gap, _rounded
from sovereign_solver_amd.problem import Problem


def _p():
    A = sp.csr_matrix(np.array([[1.0, 2.0, 0.0], [0.0, 1.0, 1.0]]))
    return Problem(c=np.array([1.0, 2.0, 3.0]), A_ub=A, b_ub=np.array([4.0, 3.0]), A_eq=sp.csr_matrix((0, 3)),
                   b_eq=np.zeros(0), lb=np.zeros(3), ub=np.array([5.0, 5.0, 5.0]),
                   integer_mask=np.array([True, True, False]))


def test_node_problem_keeps_matrices_and_drops_integrality():
    p = _p()
    q = node_problem(p, np.array([1.0, 0.0, 0.0]), np.array([2.0, 5.0, 5.0]))
    assert q.A_ub is p.A_ub and q.A_eq is p.A_eq             # one shared-matrix batch
    assert not q.integer_mask.any() and q.lb[0] == 1.0 and q.ub[0] == 2.0
    d = child_problem(p, p.lb, p.ub, 1, "down", 1.4)
    u = child_problem(p, p.lb, p.ub, 1, "up", 1.4)
    assert d.ub[1] == 1.0 and u.lb[1] == 2.0 and d.A_ub is p.A_ub


def test_rounding_and_feasibility():
    p = _p()
    xr = _rounded(p, np.array([0.9999, 1.0002, 0.5]))
    assert np.array_equal(xr, [1.0, 1.0, 0.5])                # integer entries rounded, the continuous kept
    assert _feasible(p, xr, 1e-6)
    assert not _feasible(p, np.array([3.0, 1.0, 0.0]), 1e-6)  # row 0: 5 > 4
    assert not _feasible(p, np.array([0.0, 0.0, 6.0]), 1e-6)  # bound


def test_gap_formula():
    # branch_and_bound.py:147-151: |incumbent - bound| / (1 + |incumbent|)
    assert _gap(10.0, 9.0) == 1.0 / 11.0
    assert _gap(np.inf, 1.0) == np.inf
