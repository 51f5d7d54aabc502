"""GPU: PDHG on a diagonally scaled box form (needs an MI355X).

The engine iterates on D_r K D_c and evaluates the KKT yardstick of the
original LP; the oracle's yardstick on the returned (unscaled) x, y checks
that, and the optima are the published LP optima (brandy) or HiGHS's.
"""

import numpy as np
import pytest
from scipy.optimize import linprog

from oracle import yardstick
from sovereign_solver_amd import solve_lp_pdhg
from sovereign_solver_amd.workloads import random_sparse_lp
from tests.golden.stress_instances import golden_problem, load_published

pytestmark = pytest.mark.gpu
PUB = load_published()


def _highs_obj(p):
    r = linprog(p.c, A_ub=p.A_ub if p.n_ineq else None, b_ub=p.b_ub if p.n_ineq else None,
                A_eq=p.A_eq if p.n_eq else None, b_eq=p.b_eq if p.n_eq else None,
                bounds=list(zip(p.lb, [None if np.isinf(u) else u for u in p.ub])), method="highs-ds")
    assert r.status == 0
    return r.fun


@pytest.mark.parametrize("method", ["ruiz", "geometric", "pock_chambolle", "ruiz_pc"])
def test_scaled_random_lp_fp64(gpu, method):
    p = random_sparse_lp(200, 500, 0.02, seed=5)
    r = solve_lp_pdhg(p, tol=1e-6, precision="fp64", scaling=method)
    assert r.status == "optimal"
    err = yardstick.rel_kkt(p, r.x, r.y)[0]
    assert err <= 1.02e-6
    # the device evaluated the same original-space KKT on the same point
    np.testing.assert_allclose(r.rel_kkt, err, rtol=1e-6)
    ref = _highs_obj(p)
    assert abs(r.objective - ref) <= 1e-5 * (1 + abs(ref))


def test_brandy_ruiz_fewer_iterations(gpu):
    # Netlib brandy: badly scaled rows; the reference's PDHG stops at its 25,000-iteration limit
    p = golden_problem("brandy")
    opt = float(PUB[("brandy", "lp_optimum")]["objective"])
    plain = solve_lp_pdhg(p, max_iters=100_000, tol=1e-4)
    scaled = solve_lp_pdhg(p, max_iters=100_000, tol=1e-4, scaling="ruiz")
    assert scaled.status == "optimal"
    assert yardstick.rel_kkt(p, scaled.x, scaled.y)[0] <= 1.05e-4
    assert abs(scaled.objective - opt) <= 1e-3 * abs(opt)
    assert scaled.iterations < 0.7 * plain.iterations


def test_scaled_warm_start_from_solution(gpu):
    p = random_sparse_lp(150, 375, 0.02, seed=8)
    r = solve_lp_pdhg(p, tol=1e-4, precision="fp64", scaling="ruiz")
    assert r.status == "optimal"
    # x0, y0 are given in the original space and mapped into the scaled one
    w = solve_lp_pdhg(p, tol=1e-4, precision="fp64", scaling="ruiz", x0=r.x, y0=r.y, step_scale=r.step_scale)
    assert w.status == "optimal" and w.iterations <= 100 < r.iterations
