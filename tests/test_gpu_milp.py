"""Batched branch-and-bound against HiGHS MILP (needs an MI355X).

Small MILPs — 0/1 knapsack (maximisation), set covering, a general-integer
problem with equality rows, an integer-infeasible one — are solved by
``solve_milp_batched`` (node relaxations in batches on the family engine) and
by ``scipy.optimize.milp`` (HiGHS branch-and-cut, FP64): the objectives agree
to the search's relative gap and the returned x is integral and feasible.
"""

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import Bounds, LinearConstraint, milp

from sovereign_solver_amd.milp import solve_milp_batched
from sovereign_solver_amd.problem import Problem

pytestmark = pytest.mark.gpu


def _highs(p: Problem):
    cons = []
    if p.n_ineq:
        cons.append(LinearConstraint(p.A_ub, -np.inf, p.b_ub))
    if p.n_eq:
        cons.append(LinearConstraint(p.A_eq, p.b_eq, p.b_eq))
    r = milp(p.c, constraints=cons, integrality=p.integer_mask.astype(int), bounds=Bounds(p.lb, p.ub))
    return r


def _check(p: Problem, res, ref):
    assert res.status == "optimal", res
    ref_obj = -ref.fun if p.sense == "max" else ref.fun
    assert abs(res.objective - ref_obj) <= 2e-4 * (1 + abs(ref_obj)), (res.objective, ref_obj)
    x = res.x
    m = p.integer_mask
    assert np.all(np.abs(x[m] - np.round(x[m])) <= 1e-9)
    if p.n_ineq:
        assert np.all(p.A_ub @ x <= p.b_ub + 1e-6 * (1 + np.abs(p.b_ub)))
    if p.n_eq:
        assert np.all(np.abs(p.A_eq @ x - p.b_eq) <= 1e-3 * (1 + np.abs(p.b_eq)))
    assert np.all(x >= p.lb - 1e-9) and np.all(x <= p.ub + 1e-9)


def _empty(n):
    return sp.csr_matrix((0, n)), np.zeros(0)


def test_knapsack_maximisation(gpu):
    rng = np.random.default_rng(3)
    n = 24
    w = rng.integers(5, 40, n).astype(float)
    v = w + rng.integers(0, 15, n)
    A_eq, b_eq = _empty(n)
    # sense 'max': c holds the negated values (problem.py:50-52)
    p = Problem(c=-v, A_ub=sp.csr_matrix(w[None, :]), b_ub=np.array([0.4 * w.sum()]), A_eq=A_eq, b_eq=b_eq,
                lb=np.zeros(n), ub=np.ones(n), integer_mask=np.ones(n, bool), sense="max")
    res = solve_milp_batched(p, batch=64)
    _check(p, res, _highs(p))
    assert res.nodes > 1 and res.lp_relaxation >= res.objective - 1e-6


def test_set_covering(gpu):
    rng = np.random.default_rng(11)
    m, n = 30, 60
    A = (rng.random((m, n)) < 0.12).astype(float)
    A[np.arange(m), rng.integers(0, n, m)] = 1.0            # every row coverable
    A_eq, b_eq = _empty(n)
    p = Problem(c=rng.integers(1, 20, n).astype(float), A_ub=sp.csr_matrix(-A), b_ub=-np.ones(m), A_eq=A_eq,
                b_eq=b_eq, lb=np.zeros(n), ub=np.ones(n), integer_mask=np.ones(n, bool))
    res = solve_milp_batched(p, batch=128)
    _check(p, res, _highs(p))


def test_general_integers_with_equalities(gpu):
    rng = np.random.default_rng(5)
    n, m_ub, m_eq = 16, 6, 2
    A_ub = rng.integers(0, 6, (m_ub, n)).astype(float)
    x0 = rng.integers(0, 4, n).astype(float)
    A_eq = rng.integers(0, 3, (m_eq, n)).astype(float)
    mask = np.zeros(n, bool)
    mask[: n // 2] = True                                  # mixed: half integer, half continuous
    p = Problem(c=rng.uniform(-5, 5, n), A_ub=sp.csr_matrix(A_ub), b_ub=A_ub @ x0 + 3, A_eq=sp.csr_matrix(A_eq),
                b_eq=A_eq @ x0, lb=np.zeros(n), ub=np.full(n, 6.0), integer_mask=mask)
    res = solve_milp_batched(p, batch=64)
    _check(p, res, _highs(p))


def test_integer_infeasible(gpu):
    # x1 + x2 = 1 and x1 - x2 = 0.5 with x integer: the relaxation is feasible (0.75, 0.25), no integer point is
    p = Problem(c=np.array([1.0, 1.0]), A_ub=sp.csr_matrix((0, 2)), b_ub=np.zeros(0),
                A_eq=sp.csr_matrix(np.array([[1.0, 1.0], [1.0, -1.0]])), b_eq=np.array([1.0, 0.5]),
                lb=np.zeros(2), ub=np.full(2, 3.0), integer_mask=np.ones(2, bool))
    res = solve_milp_batched(p, batch=16, node_limit=200)
    assert res.status == "infeasible" and res.x is None
