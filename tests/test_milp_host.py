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


def _stub_solver(status, x_of):
    """A stand-in for the batched LP call: every node LP ends with ``status`` at the
    point ``x_of(problem)``; records the iteration limits it was called with."""
    from sovereign_solver_amd.result import SolveResult
    calls = []

    def solve(problems, warm_start=None, tol=1e-4, max_iters=0, engine="auto"):
        calls.append((len(problems), max_iters))
        return [SolveResult(status, x_of(q), float(np.dot(q.c, np.nan_to_num(x_of(q)))), max_iters,
                            y=np.zeros(q.n_constraints), step_scale=1.0) for q in problems]
    return solve, calls


def test_unsolved_integral_nodes_are_retried_and_never_reported_optimal(monkeypatch):
    import sovereign_solver_amd.milp.branch_and_bound as BB
    # iteration limit at an integral point: the rounded point becomes the incumbent, but the
    # relaxation is not solved, so the subtree is not proven and the search must not say optimal
    solve, calls = _stub_solver("iteration_limit", lambda q: np.clip(np.array([1.0, 1.0, 0.0]), q.lb, q.ub))
    monkeypatch.setattr(BB, "solve_lp_pdhg_batch", solve)
    r = BB.solve_milp_batched(_p(), max_lp_iters=100, node_limit=50)
    assert r.status == "node_limit" and r.x is not None
    assert [it for _, it in calls] == [100, 400, 1600]          # root, then two retries with 4x the limit


def test_nan_points_end_the_search_without_looping(monkeypatch):
    import sovereign_solver_amd.milp.branch_and_bound as BB
    solve, calls = _stub_solver("numerical_error", lambda q: np.full(q.n_vars, np.nan))
    monkeypatch.setattr(BB, "solve_lp_pdhg_batch", solve)
    r = BB.solve_milp_batched(_p(), max_lp_iters=100, node_limit=1000)
    assert r.status == "node_limit" and r.x is None
    assert len(calls) == 3 and r.nodes == 3                     # the root and its two retries, then stop


def test_unsolved_fractional_nodes_still_branch(monkeypatch):
    import sovereign_solver_amd.milp.branch_and_bound as BB
    # every node stops at the iteration limit at a fractional point: the search branches on it
    # (the bound stays the parent's) until the bounds fix the variables, and ends unproven
    solve, calls = _stub_solver("iteration_limit", lambda q: np.clip(np.array([0.5, 0.5, 0.0]), q.lb, q.ub))
    monkeypatch.setattr(BB, "solve_lp_pdhg_batch", solve)
    r = BB.solve_milp_batched(_p(), max_lp_iters=100, node_limit=200)
    assert r.status == "node_limit"
    assert sum(n for n, _ in calls) > 1
