"""Relative KKT yardstick on the reference-form lift.  TEST INFRASTRUCTURE.

Every GPU result, whatever form the engine ran in, is judged by the
reference KKT formula (``solver/lp/pdhg.py:184-196``) evaluated on the
reference standard form (``pdhg.py:62-125``).  This module builds that lift
*explicitly* — the standard-form matrix from :func:`oracle.pdhg_ref.to_standard_form`,
a standard-form primal ``z`` and dual ``y_ref`` — and applies the formula, so
it shares no code with the device's fused evaluation.

Lift of a box-form pair (x in the original space, y = duals of
``[A_ub; A_eq]`` with y ≤ 0 on the ``<=`` rows):

* ``z = [x − lb_f ; max(b_ub − A_ub x, 0) ; (ub − x)_F]``
* ``r = c − A_ubᵀ y_ub − A_eqᵀ y_eq``; upper-bound-row duals ``w = min(r_F, 0)``
* ``y_ref = row_sign · [y_ub ; w ; y_eq]``

Pinned in ``tests/test_oracle.py``: on problems whose box form has no
bounds or inequality rows the lift is the identity and the value must equal
``pdhg_ref.kkt`` on the raw data; at an optimal vertex from HiGHS it must be
~0; hand-computed 1-variable cases.
"""

from __future__ import annotations

import numpy as np

from .pdhg_ref import kkt, to_standard_form


def lift(p, x, y):
    sf = to_standard_form(p)
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    m_ub = p.n_ineq
    y_ub, y_eq = y[:m_ub], y[m_ub:]
    s1 = np.maximum(np.asarray(p.b_ub, float) - p.A_ub @ x, 0.0)
    F = sf.ub_cols
    s2 = (np.asarray(p.ub, float) - x)[F]
    z = np.concatenate([x - sf.lb_shift, s1, s2])
    r = np.asarray(p.c, float) - p.A_ub.T @ y_ub - p.A_eq.T @ y_eq
    w = np.minimum(r[F], 0.0)
    y_ref = np.concatenate([y_ub, w, y_eq]) * sf.row_sign
    return sf, z, y_ref


def rel_kkt(p, x, y):
    """(max, rp, rd, gap) of the box-form pair (x, y) by the reference formula."""
    sf, z, yr = lift(p, x, y)
    A = sf.A
    norm_b = max(1.0, float(np.linalg.norm(sf.b)))
    norm_c = max(1.0, float(np.linalg.norm(sf.c)))
    return kkt(A, A.T.tocsr(), sf.b, sf.c, norm_b, norm_c, z, yr)


def rel_kkt_standard(p, z, y_std):
    """(max, rp, rd, gap) of a standard-form pair, exactly as pdhg.py evaluates it."""
    sf = to_standard_form(p)
    norm_b = max(1.0, float(np.linalg.norm(sf.b)))
    norm_c = max(1.0, float(np.linalg.norm(sf.c)))
    return kkt(sf.A, sf.A.T.tocsr(), sf.b, sf.c, norm_b, norm_c, np.asarray(z, float), np.asarray(y_std, float))
