"""Farkas certificates of primal infeasibility.  TEST INFRASTRUCTURE.

For a box-form LP (rows ``A_ub x <= b_ub``, ``A_eq x = b_eq``, bounds
``lb <= x <= ub``) and a dual vector y ordered ``[A_ub rows; A_eq rows]``
with y <= 0 on the ``<=`` rows, every feasible x satisfies
``y'K x >= y'b``; with ``z = x - lb_f`` (lb_f = finite lower bound, else 0)
that is ``(K'y)'z >= (b - K lb_f)'y``.  So

    F(y) = (b - K lb_f)'y - max over the shifted box of (K'y)'z  > 0

proves that no feasible x exists.  The maximum is taken column by column:
``g_j * (ub_j - lb_f_j)`` for ``g_j > 0`` (+inf without an upper bound),
+inf for ``g_j < 0`` without a lower bound, else 0.  The reference's PDHG
has no infeasibility detection, so this is pinned by the lemma itself
(``tests/test_oracle.py``: a hand-made certificate, and F <= 0 for random y
on feasible LPs), and on the GPU by HiGHS declaring the same LPs infeasible.

Rays of decrease (dual infeasibility): d with ``c'd < 0``, ``A_eq d = 0``,
``A_ub d <= 0``, ``d_j >= 0`` where lb_j is finite and ``d_j <= 0`` where ub_j
is finite.  From a feasible x, ``x + t d`` stays feasible for every t >= 0
and its objective falls without bound.  ``ray_violations`` measures how far
a given d is from that cone; it is pinned by hand-made rays and by the
lemma that no such d exists for a bounded LP (``tests/test_oracle.py``).
"""

import numpy as np
import scipy.sparse as sp


def farkas_value(p, y) -> float:
    """F(y) on the host in FP64 with the original data (y is clipped to y <= 0
    on the ``<=`` rows first, as the engine's projection keeps it)."""
    K = sp.vstack([sp.csr_matrix(p.A_ub), sp.csr_matrix(p.A_eq)]).tocsr()
    b = np.concatenate([np.asarray(p.b_ub, float), np.asarray(p.b_eq, float)])
    y = np.asarray(y, float).copy()
    y[: p.n_ineq] = np.minimum(y[: p.n_ineq], 0.0)
    lb = np.asarray(p.lb, float)
    ub = np.asarray(p.ub, float)
    lbf = np.where(np.isfinite(lb), lb, 0.0)
    g = K.T @ y
    span = ub - lbf
    sup = np.where(g > 0, np.where(np.isfinite(span), g * span, np.inf),
                   np.where((g < 0) & ~np.isfinite(lb), np.inf, 0.0))
    return float((b - K @ lbf) @ y - sup.sum())


def ray_violations(p, d):
    """``(c'd, |viol(K d)|_2, |viol_bounds(d)|_2)`` on the host in FP64 with the
    original data: the rows' violations are ``A_eq d`` and ``max(A_ub d, 0)``,
    the bounds' ``min(d_j, 0)`` where lb_j is finite and ``max(d_j, 0)`` where
    ub_j is finite."""
    d = np.asarray(d, float)
    ru = sp.csr_matrix(p.A_ub) @ d
    re = sp.csr_matrix(p.A_eq) @ d
    rows = np.concatenate([np.maximum(ru, 0.0), re])
    lb = np.asarray(p.lb, float)
    ub = np.asarray(p.ub, float)
    bnd = np.where(np.isfinite(lb) & (d < 0), d, 0.0) + np.where(np.isfinite(ub) & (d > 0), d, 0.0)
    return float(np.asarray(p.c, float) @ d), float(np.linalg.norm(rows)), float(np.linalg.norm(bnd))
