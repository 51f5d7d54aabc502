"""The host's FP64 ray certificate (``lp.pdhg.ray_is_certified``), no GPU needed.

Pinned by hand-made rays and by the lemma behind it: for a bounded LP with a dual
solution (y*, r*), c'd >= -|y*| |viol(K d)| - |r*| |viol_bounds(d)| for every d, so
a bounded LP's directions fail the test while its dual norm is below 1/eps."""

import numpy as np
import scipy.sparse as sp

from sovereign_solver_amd.lp.pdhg import RAY_EPS_HOST, ray_is_certified, ray_quality
from sovereign_solver_amd.problem import Problem


def _lp(c, A_ub=None, b_ub=None, A_eq=None, b_eq=None, lb=None, ub=None):
    n = len(c)
    emp = sp.csr_matrix((0, n))
    return Problem(c=np.asarray(c, float), A_ub=emp if A_ub is None else sp.csr_matrix(np.asarray(A_ub, float)),
                   b_ub=np.zeros(0) if b_ub is None else np.asarray(b_ub, float),
                   A_eq=emp if A_eq is None else sp.csr_matrix(np.asarray(A_eq, float)),
                   b_eq=np.zeros(0) if b_eq is None else np.asarray(b_eq, float),
                   lb=np.zeros(n) if lb is None else np.asarray(lb, float),
                   ub=np.full(n, np.inf) if ub is None else np.asarray(ub, float), integer_mask=np.zeros(n, bool))


def test_exact_ray_is_certified():
    # min -x1 + 0.5 x2 s.t. x1 - x2 = 0, -x1 <= 0: d = (1, 1) lowers c'x by 0.5 per unit
    p = _lp([-1.0, 0.5], A_ub=[[-1.0, 0.0]], b_ub=[0.0], A_eq=[[1.0, -1.0]], b_eq=[0.0])
    cd, vr, vb = ray_quality(p, np.array([1.0, 1.0]))
    assert (cd, vr, vb) == (-0.5, 0.0, 0.0)
    assert ray_is_certified(p, np.array([1.0, 1.0]))
    assert not ray_is_certified(p, np.array([-1.0, -1.0]))          # c'd > 0, and the bounds break


def test_bound_violation_counts():
    # min -x with x <= 5: d = 1 leaves the box on its finite side
    p = _lp([-1.0], ub=[5.0])
    cd, vr, vb = ray_quality(p, np.array([1.0]))
    assert cd == -1.0 and vr == 0.0 and vb == 1.0
    assert not ray_is_certified(p, np.array([1.0]))


def test_far_optimum_directions_fail_below_the_dual_bound():
    # min -x1 s.t. x1 = x2, a x2 <= 1: bounded, |y*| = 1/a; along (1, 1) the normalised
    # row violation is a, so the direction passes only when 1/a >= 1/eps
    for a, passes in ((1e-5, False), (1e-4, False), (3e-4, False), (1e-7, True)):
        p = _lp([-1.0, 0.0], A_ub=[[0.0, a]], b_ub=[1.0], A_eq=[[1.0, -1.0]], b_eq=[0.0])
        d = np.array([1.0, 1.0])
        cd, vr, vb = ray_quality(p, d)
        assert np.isclose(vr / -cd, a)
        assert ray_is_certified(p, d) == passes, a
    assert RAY_EPS_HOST == 2e-6


def test_random_directions_of_bounded_lps_fail():
    # the lemma: with HiGHS duals of a bounded LP no direction beats -|y*| vr - |r*| vb
    from scipy.optimize import linprog
    rng = np.random.default_rng(0)
    for _ in range(5):
        A = rng.standard_normal((6, 10))
        x0 = rng.uniform(0, 1, 10)
        p = _lp(A.T @ rng.standard_normal(6) + np.abs(rng.standard_normal(10)), A_eq=A, b_eq=A @ x0)
        h = linprog(p.c, A_eq=A, b_eq=p.b_eq, bounds=(0, None), method="highs-ds")
        y, r = h.eqlin.marginals, h.lower.marginals
        for _ in range(50):
            d = rng.standard_normal(10)
            cd, vr, vb = ray_quality(p, d)
            assert cd >= -np.linalg.norm(y) * vr - np.linalg.norm(r) * vb - 1e-9
            assert not ray_is_certified(p, d)
