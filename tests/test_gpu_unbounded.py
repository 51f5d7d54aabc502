"""GPU: dual infeasibility (unboundedness) in the single-instance engine (needs an MI355X).

An LP with a ray of decrease ends with ``status="unbounded"``, ``x=None``,
objective ``-inf`` and the ray ``d`` (``max |d_j| = 1``), which the host
checks in FP64 with the original data (``oracle.farkas.ray_violations``):
``c'd < 0`` and row plus bound violations below 2e-6 of ``-c'd``.  HiGHS
reports the same LPs unbounded.  Bounded LPs with free columns, and bounded
LPs whose optimum lies far out along a near-ray (duals up to 1e5), the
classes a loose test would flag, do not end unbounded.
"""

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import linprog

from oracle.farkas import ray_violations
from sovereign_solver_amd import solve_lp_pdhg
from sovereign_solver_amd.problem import Problem
from sovereign_solver_amd.workloads import random_sparse_lp

pytestmark = pytest.mark.gpu


def _highs_status(p):
    r = linprog(p.c, A_ub=p.A_ub if p.n_ineq else None, b_ub=p.b_ub if p.n_ineq else None,
                A_eq=p.A_eq if p.n_eq else None, b_eq=p.b_eq if p.n_eq else None,
                bounds=list(zip([None if np.isinf(v) else v for v in p.lb],
                                [None if np.isinf(v) else v for v in p.ub])), method="highs")
    return r.status


def _free_columns(seed):
    """30 '=' rows over 60 columns, every third free below and every second free above."""
    p = random_sparse_lp(30, 60, 0.1, seed=seed, family=2)
    lb = np.where(np.arange(60) % 3 == 0, -np.inf, p.lb)
    ub = np.where(np.arange(60) % 2 == 0, np.inf, p.ub)
    return Problem(c=p.c, A_ub=p.A_ub, b_ub=p.b_ub, A_eq=p.A_eq, b_eq=p.b_eq, lb=lb, ub=ub,
                   integer_mask=p.integer_mask, name=f"free_columns_{seed}")


def _linked_pair():
    """A bounded random LP plus columns u, v >= 0 with costs -1 and 0.5 that cancel in one
    '=' row (u - v) and loosen one '<=' row (-u): d = e_u + e_v lowers c'x by 0.5 per unit."""
    base = random_sparse_lp(60, 150, 0.05, seed=4)
    n = base.n_vars
    A_eq = sp.hstack([base.A_eq, sp.csr_matrix(([1.0, -1.0], ([3, 3], [0, 1])), shape=(base.n_eq, 2))]).tocsr()
    A_ub = sp.vstack([sp.hstack([base.A_ub, sp.csr_matrix((base.n_ineq, 2))]),
                      sp.csr_matrix(([-1.0], ([0], [n])), shape=(1, n + 2))]).tocsr()
    b_ub = np.concatenate([base.b_ub, [0.0]])
    return Problem(c=np.concatenate([base.c, [-1.0, 0.5]]), A_ub=A_ub, b_ub=b_ub, A_eq=A_eq, b_eq=base.b_eq,
                   lb=np.concatenate([base.lb, [0.0, 0.0]]), ub=np.concatenate([base.ub, [np.inf, np.inf]]),
                   integer_mask=np.zeros(n + 2, bool), name="linked_pair")


def _check_ray(p, r):
    assert r.status == "unbounded", (r.status, r.iterations)
    assert r.x is None and r.objective == -np.inf
    assert np.abs(r.ray).max() == pytest.approx(1.0)
    cd, vr, vb = ray_violations(p, r.ray)
    assert cd < 0.0
    # the engine's own bound (1e-6 on the device, 2e-6 on the host in FP64)
    assert vr + vb <= 2e-6 * -cd, (cd, vr, vb)


def _far_optimum(a):
    """min -x1 s.t. x1 - x2 = 0, a x2 <= 1, x >= 0: bounded, optimum -1/a at x = (1/a, 1/a),
    every dual solution has |y| >= 1/a.  Early iterates move along (1, 1), whose row
    violation a x2 is small against c'd = -x1 without being zero."""
    return Problem(c=np.array([-1.0, 0.0]), A_ub=sp.csr_matrix(np.array([[0.0, a]])), b_ub=np.array([1.0]),
                   A_eq=sp.csr_matrix(np.array([[1.0, -1.0]])), b_eq=np.array([0.0]), lb=np.zeros(2),
                   ub=np.full(2, np.inf), integer_mask=np.zeros(2, bool), name=f"far_optimum_{a:g}")


@pytest.mark.parametrize("make", [lambda: _free_columns(1), lambda: _free_columns(3), _linked_pair],
                         ids=["free_columns_1", "free_columns_3", "linked_pair"])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_unbounded_lp_returns_ray(gpu, make, precision):
    p = make()
    assert _highs_status(p) == 3                         # HiGHS: unbounded
    r = solve_lp_pdhg(p, max_iters=100_000, tol=1e-4, precision=precision)
    assert r.iterations < 100_000
    _check_ray(p, r)


def test_unbounded_lp_with_scaling(gpu):
    p = _linked_pair()
    r = solve_lp_pdhg(p, max_iters=100_000, tol=1e-4, scaling="ruiz_pc")
    _check_ray(p, r)                                     # the ray comes back in the original space


def test_reference_scheme_runs_to_the_limit(gpu):
    # the reference's PDHG has no unboundedness detection
    r = solve_lp_pdhg(_linked_pair(), max_iters=2_000, tol=1e-4, scheme="reference")
    assert r.status == "iteration_limit" and r.iterations == 2_000


@pytest.mark.parametrize("seed", [0, 2, 5])
def test_bounded_lps_with_free_columns_are_not_flagged(gpu, seed):
    p = _free_columns(seed)
    assert _highs_status(p) == 0
    for precision in ("fp32", "fp64"):
        r = solve_lp_pdhg(p, max_iters=100_000, tol=1e-4, precision=precision)
        assert r.status == "optimal", (precision, r.status, r.iterations)


@pytest.mark.parametrize("a", [1e-5, 1e-4, 3e-4])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_far_optimum_is_not_flagged_unbounded(gpu, a, precision):
    # a bounded LP whose duals are large (1/a) but below the 1e6 the ray test admits
    p = _far_optimum(a)
    assert _highs_status(p) == 0
    r = solve_lp_pdhg(p, max_iters=100_000, tol=1e-4, precision=precision)
    assert r.status != "unbounded" and r.x is not None, (r.status, r.iterations)
    if a >= 1e-4:                        # the nearer optima are reached within the limit
        assert r.status == "optimal", (r.status, r.iterations)
        assert abs(r.objective + 1.0 / a) <= 1e-4 * (1.0 + 1.0 / a), r.objective
