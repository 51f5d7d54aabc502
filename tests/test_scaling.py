"""Diagonal scaling of the box form (CPU): the scalers against closed forms and
invariants, and the scaled form's equivalence to the original LP."""

import numpy as np
import pytest
import scipy.sparse as sp

from sovereign_solver_amd.engine import scaling as S
from sovereign_solver_amd.engine.forms import box_form, extract, make_form
from sovereign_solver_amd.problem import Problem
from sovereign_solver_amd.workloads import random_sparse_lp


def _wide_range(m, n, density, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density, random_state=seed, format="csr")
    A.data = np.sign(rng.standard_normal(A.nnz)) * np.exp(3.0 * rng.standard_normal(A.nnz))
    return A


def test_ruiz_diagonal_closed_form():
    # one pass divides each line by sqrt(max |entry|): diag(4, 1/9) -> identity
    Dr, Dc = S.ruiz(sp.csr_matrix(np.diag([4.0, 1.0 / 9.0])))
    np.testing.assert_allclose(Dr, [0.5, 3.0])
    np.testing.assert_allclose(Dc, [0.5, 3.0])


def test_ruiz_bounds_entries_and_equilibrates():
    A = _wide_range(60, 90, 0.08, 1)
    Dr, Dc = S.ruiz(A)
    B = S.scale_matrix(A, Dr, Dc)
    # every pass leaves |entries| <= 1 (|a_ij| <= max of its row and of its column)
    assert np.abs(B.data).max() <= 1.0 + 1e-12

    def deviation(M):
        r = np.abs(M).max(axis=1).toarray().ravel()
        c = np.abs(M).max(axis=0).toarray().ravel()
        return max(np.abs(r[r > 0] - 1).max(), np.abs(c[c > 0] - 1).max())

    assert deviation(B) < 0.5 * deviation(A)
    # the pattern and entry order are kept
    np.testing.assert_array_equal(B.indptr, A.indptr)
    np.testing.assert_array_equal(B.indices, A.indices)


def test_ruiz_empty_lines_keep_unit_scale():
    A = sp.csr_matrix(np.array([[2.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 8.0, 0.0]]))
    Dr, Dc = S.ruiz(A)
    assert Dr[1] == 1.0 and Dc[2] == 1.0
    np.testing.assert_allclose(np.abs(S.scale_matrix(A, Dr, Dc).data), 1.0)


def test_geometric_rank_one_halves_the_log_spread_per_pass():
    # for a_ij = u_i v_j each row pass halves the spread of log u, each column
    # pass that of log v, so five passes divide the spread of log|a| by 2^5
    rng = np.random.default_rng(0)
    u, v = np.exp(3 * rng.standard_normal(20)), np.exp(3 * rng.standard_normal(30))
    A = sp.csr_matrix(np.outer(u, v))
    B = S.scale_matrix(A, *S.geometric(A))
    np.testing.assert_allclose(np.ptp(np.log(B.data)), np.ptp(np.log(A.data)) / 32, rtol=1e-9)


def test_pock_chambolle_norm_bound():
    # Schur test: || D_r K D_c ||_2 <= 1 with D = 1/sqrt(absolute line sums); equality for all-ones
    A = _wide_range(40, 70, 0.1, 2)
    B = S.scale_matrix(A, *S.pock_chambolle(A)).toarray()
    assert np.linalg.norm(B, 2) <= 1.0 + 1e-12
    J = sp.csr_matrix(np.ones((5, 7)))
    np.testing.assert_allclose(np.linalg.norm(S.scale_matrix(J, *S.pock_chambolle(J)).toarray(), 2), 1.0)


@pytest.mark.parametrize("method", ["ruiz", "geometric", "pock_chambolle", "ruiz_pc"])
def test_scaled_box_form_is_the_same_lp(method):
    rng = np.random.default_rng(3)
    p = random_sparse_lp(40, 100, 0.05, seed=3)
    A_ub = _wide_range(15, 100, 0.1, 4)
    p = Problem(c=p.c * np.exp(rng.standard_normal(100)), A_ub=A_ub, b_ub=rng.uniform(1, 5, 15),
                A_eq=p.A_eq, b_eq=p.b_eq, lb=np.where(rng.random(100) < 0.3, -2.0, 0.0), ub=p.ub,
                integer_mask=p.integer_mask)
    f0, f = box_form(p), box_form(p, method)
    Dr, Dc = f.row_scale, f.col_scale
    x = rng.uniform(p.lb, p.ub)
    xs = x / Dc
    np.testing.assert_allclose(f.K @ xs - f.b, Dr * (f0.K @ x - f0.b), rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(f.c @ xs, p.c @ x, rtol=1e-12)
    np.testing.assert_allclose(f.lo * Dc, p.lb, rtol=1e-14)
    np.testing.assert_allclose(f.hi * Dc, p.ub, rtol=1e-14)
    np.testing.assert_allclose(f.b_shift, Dr * f0.b_shift, rtol=1e-12)
    # the yardstick's normalisers are the original LP's
    assert f.norm_b == f0.norm_b and f.norm_c == f0.norm_c
    # extract maps the scaled point back
    y = rng.standard_normal(f.m)
    x_out, obj, y_out = extract(f, p, xs, y / Dr)
    np.testing.assert_allclose(x_out, x, rtol=1e-12)
    np.testing.assert_allclose(y_out, y, rtol=1e-12)
    np.testing.assert_allclose(obj, p.c @ x, rtol=1e-12)


def test_scaling_arguments():
    p = random_sparse_lp(10, 20, 0.2, seed=0)
    with pytest.raises(ValueError):
        box_form(p, "bogus")
    with pytest.raises(ValueError):
        make_form(p, "reference", "ruiz")
    f = box_form(p, "none")
    assert f.row_scale is None and f.col_scale is None
