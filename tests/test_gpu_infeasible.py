"""GPU: primal infeasibility in the single-instance engine (needs an MI355X).

An infeasible LP ends with ``status="infeasible"`` and its dual ``y`` is a
Farkas certificate, checked on the host in FP64 with the original data
(``oracle.farkas``); HiGHS agrees that the LP is infeasible.  The reference
scheme keeps the reference's behaviour (no detection).
"""

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import linprog

from oracle.farkas import farkas_value, ray_violations
from sovereign_solver_amd import solve_lp_pdhg
from sovereign_solver_amd.problem import Problem
from sovereign_solver_amd.workloads import SetCoverNodes, random_sparse_lp

pytestmark = pytest.mark.gpu


def _highs_status(p):
    r = linprog(p.c, A_ub=p.A_ub if p.n_ineq else None, b_ub=p.b_ub if p.n_ineq else None,
                A_eq=p.A_eq if p.n_eq else None, b_eq=p.b_eq if p.n_eq else None,
                bounds=list(zip([None if np.isinf(l) else l for l in p.lb],
                                [None if np.isinf(u) else u for u in p.ub])), method="highs")
    return r.status


def _uncovered_set_cover():
    """A set-cover node whose row 3 has every variable fixed to 0."""
    sc = SetCoverNodes(rows=100, cols=1000, density=0.01)
    p = sc.node(3)
    lb, ub = p.lb.copy(), p.ub.copy()
    cols = sc.A[3].indices
    lb[cols] = 0.0
    ub[cols] = 0.0
    return sc.node_from(lb, ub, name="uncovered")


def _equality_beyond_box():
    """Random '=' rows plus one row asking sum(x) above what the box allows."""
    base = random_sparse_lp(60, 150, 0.05, seed=4)
    n = base.n_vars
    A_eq = sp.vstack([base.A_eq, sp.csr_matrix(np.ones((1, n)))]).tocsr()
    b_eq = np.concatenate([base.b_eq, [1.5 * base.ub.sum()]])
    return Problem(c=base.c, A_ub=base.A_ub, b_ub=base.b_ub, A_eq=A_eq, b_eq=b_eq, lb=base.lb, ub=base.ub,
                   integer_mask=base.integer_mask, name="beyond_box")


def _conflict(lo=-5.0, hi=5.0):
    """Random '<=' rows and a contradicting pair (x1 + x2 <= 1, -x1 - x2 <= -3) over [lo, hi]^n."""
    rng = np.random.default_rng(2)
    n = 40
    A = sp.random(20, n, 0.2, random_state=2, format="csr")
    x0 = rng.standard_normal(n)
    rows = sp.vstack([A, sp.csr_matrix(np.r_[[1.0, 1.0], np.zeros(n - 2)][None, :]),
                      sp.csr_matrix(np.r_[[-1.0, -1.0], np.zeros(n - 2)][None, :])]).tocsr()
    b = np.concatenate([A @ x0 + 1.0, [1.0, -3.0]])
    return Problem(c=rng.standard_normal(n), A_ub=rows, b_ub=b, A_eq=sp.csr_matrix((0, n)), b_eq=np.zeros(0),
                   lb=np.full(n, lo), ub=np.full(n, hi), integer_mask=np.zeros(n, bool), name="conflict")


@pytest.mark.parametrize("make", [_uncovered_set_cover, _equality_beyond_box, _conflict])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_infeasible_lp_returns_certificate(gpu, make, precision):
    p = make()
    assert _highs_status(p) == 2                         # HiGHS: infeasible
    r = solve_lp_pdhg(p, max_iters=100_000, tol=1e-4, precision=precision)
    assert r.status == "infeasible", (r.status, r.iterations)
    assert r.iterations < 100_000
    assert farkas_value(p, r.y) > 0.0


def test_free_columns_conflict_is_not_certified(gpu):
    # with free columns a Farkas certificate needs K'y = 0 there exactly, which PDHG's
    # iterates only approach, so the exact test never fires.  This LP (40 free columns,
    # 22 rows) is dual infeasible as well: c'x has a ray of decrease, which the
    # unboundedness test reports (status "unbounded" = dual infeasible)
    p = _conflict(-np.inf, np.inf)
    assert _highs_status(p) == 2
    r = solve_lp_pdhg(p, max_iters=5_000, tol=1e-4)
    assert r.status == "unbounded"
    cd, vr, vb = ray_violations(p, r.ray)
    assert cd < 0.0 and vr <= 2e-3 * -cd and vb == 0.0


def test_infeasible_lp_with_scaling(gpu):
    p = _uncovered_set_cover()
    r = solve_lp_pdhg(p, max_iters=100_000, tol=1e-4, scaling="ruiz_pc")
    assert r.status == "infeasible"
    assert farkas_value(p, r.y) > 0.0                    # y comes back in the original space


def test_reference_scheme_keeps_reference_behaviour(gpu):
    # the reference's PDHG has no infeasibility detection: it runs to the iteration limit
    p = _uncovered_set_cover()
    r = solve_lp_pdhg(p, max_iters=2_000, tol=1e-4, scheme="reference")
    assert r.status == "iteration_limit" and r.iterations == 2_000


@pytest.mark.parametrize("seed", [0, 1])
def test_feasible_lps_are_not_flagged(gpu, seed):
    for p in (random_sparse_lp(150, 375, 0.02, seed=seed), SetCoverNodes(rows=100, cols=1000, density=0.01).node(seed)):
        r = solve_lp_pdhg(p, max_iters=100_000, tol=1e-4)
        assert r.status == "optimal"
