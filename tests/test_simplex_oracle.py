"""Pin the own FP64 simplex oracle: HiGHS on random and reference LPs,
brute-force vertex enumeration on tiny LPs (no GPU)."""

import itertools

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import linprog

from oracle import simplex, yardstick
from sovereign_solver_amd.problem import Problem
from sovereign_solver_amd.workloads import random_sparse_lp, transportation_lp
from tests.golden.stress_instances import golden_problem, load_published


def _highs(p):
    r = linprog(p.c, A_ub=p.A_ub if p.n_ineq else None, b_ub=p.b_ub if p.n_ineq else None,
                A_eq=p.A_eq if p.n_eq else None, b_eq=p.b_eq if p.n_eq else None,
                bounds=list(zip(p.lb, [None if np.isinf(u) else u for u in p.ub])), method="highs-ds")
    return r


@pytest.mark.parametrize("seed", range(4))
def test_simplex_matches_highs_random(seed):
    p = random_sparse_lp(40, 100, 0.08, seed=seed)
    s = simplex.solve(p)
    h = _highs(p)
    assert s.status == "optimal" and h.status == 0
    assert abs(s.objective - h.fun) <= 1e-8 * (1 + abs(h.fun))
    # optimal pair: reference KKT of the lift is ~0
    assert yardstick.rel_kkt(p, s.x, s.y)[0] < 1e-8


def test_simplex_matches_highs_transport():
    p = transportation_lp(6, 9, seed=2)
    s = simplex.solve(p)
    assert s.status == "optimal"
    assert abs(s.objective - _highs(p).fun) <= 1e-8 * abs(s.objective)


@pytest.mark.parametrize("name", ["afiro", "blending", "random_n50_m30"])
def test_simplex_matches_published_optima(name):
    p = golden_problem(name)
    s = simplex.solve(p)
    opt = float(load_published()[(name, "lp_optimum")]["objective"])
    assert s.status == "optimal"
    assert abs(s.objective - opt) <= 1e-7 * max(1.0, abs(opt))


def _brute_force(c, A_ub, b_ub, lb, ub):
    """Enumerate vertices of {A_ub x <= b_ub, lb <= x <= ub} in R^2 / R^3."""
    n = len(c)
    rows = [(A_ub[i], b_ub[i]) for i in range(len(b_ub))]
    for j in range(n):
        e = np.zeros(n); e[j] = 1.0
        rows.append((e, ub[j]))
        rows.append((-e, -lb[j]))
    best = np.inf
    for idx in itertools.combinations(range(len(rows)), n):
        M = np.array([rows[i][0] for i in idx])
        if abs(np.linalg.det(M)) < 1e-12:
            continue
        v = np.linalg.solve(M, np.array([rows[i][1] for i in idx]))
        if all(a @ v <= bb + 1e-9 for a, bb in rows):
            best = min(best, c @ v)
    return best


@pytest.mark.parametrize("seed", range(6))
def test_simplex_brute_force_tiny(seed):
    rng = np.random.default_rng(seed)
    n = 2 + seed % 2
    A = rng.uniform(-1, 2, (3, n))
    b = rng.uniform(1, 4, 3)
    c = rng.standard_normal(n)
    lb = np.zeros(n)
    ub = rng.uniform(1, 3, n)
    p = Problem(c=c, A_ub=sp.csr_matrix(A), b_ub=b, A_eq=sp.csr_matrix((0, n)), b_eq=np.zeros(0), lb=lb, ub=ub,
                integer_mask=np.zeros(n, bool))
    s = simplex.solve(p)
    assert s.status == "optimal"
    assert abs(s.objective - _brute_force(c, A, b, lb, ub)) <= 1e-9


def test_simplex_detects_infeasible_and_unbounded():
    inf = Problem(c=np.array([1.0]), A_ub=sp.csr_matrix([[1.0]]), b_ub=np.array([-1.0]), A_eq=sp.csr_matrix((0, 1)),
                  b_eq=np.zeros(0), lb=np.zeros(1), ub=np.full(1, np.inf), integer_mask=np.zeros(1, bool))
    assert simplex.solve(inf).status == "infeasible"
    unb = Problem(c=np.array([-1.0, 0.0]), A_ub=sp.csr_matrix([[1.0, -1.0]]), b_ub=np.array([1.0]),
                  A_eq=sp.csr_matrix((0, 2)), b_eq=np.zeros(0), lb=np.zeros(2), ub=np.full(2, np.inf),
                  integer_mask=np.zeros(2, bool))
    assert simplex.solve(unb).status == "unbounded"
