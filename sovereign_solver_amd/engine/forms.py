"""Host-side formulations of a :class:`Problem` for the device engine.

Two forms feed the same kernels:

``box`` (the engine's native form)
    K = [A_ub; A_eq] with row kinds (``<=`` rows keep a dual y ≤ 0, ``=`` rows a
    free dual), ``lb ≤ x ≤ ub`` enforced by projection.  No slack columns, no
    upper-bound rows, no shift.

``reference`` (parity mode)
    The equality standard form of ``solver/lp/pdhg.py:62-125``: rows
    ``[A_ub; E_F; A_eq]`` (E_F = one row per finite upper bound), columns
    ``[x - lb_f; slacks of the <= rows]``, rows with ``b_i < -1e-6`` negated,
    every row ``=``, ``x ≥ 0``.  Running it reproduces the reference trajectory.

Both carry the FP64 data of the KKT *yardstick*: the reference formula
(``pdhg.py:184-196``) evaluated on the reference-form lift of the current
primal/dual pair.  For the box form the lift is (F = columns with a finite
upper bound, lb_f = finite lower bound else 0, r = c − Kᵀy):

* primal residual rows: ``max(K_i x − b_i, 0)`` on ``<=`` rows (slack =
  ``max(b − Kx, 0)``), ``K_i x − b_i`` on ``=`` rows; upper-bound rows are
  satisfied exactly (slack ``u − x``);
* upper-bound-row duals ``w_j = min(r_j, 0)`` for j ∈ F, so the dual residual
  is ``min(r_j, 0)`` on columns outside F and ``max(y_i, 0)`` on ``<=`` rows;
* primal objective ``cᵀ(x − lb_f)``, dual objective
  ``(b − K lb_f)ᵀy + Σ_F (u_j − lb_f,j) w_j``;
* normalisers ``max(1, ‖b_ref‖)`` and ``max(1, ‖c‖)`` with
  ``b_ref = [b − K lb_f ; (u − lb_f)_F]``.

A box form may be *scaled* (``box_form(p, scaling=...)``, methods in
:mod:`.scaling`): K, b, c and the bounds are those of ``D_r K D_c``, and
``row_scale`` / ``col_scale`` carry ``D_r`` / ``D_c`` so that the kernels
evaluate the yardstick on the original LP; ``norm_b`` and ``norm_c`` are the
original LP's.  :func:`extract` maps the solution back.
"""

from __future__ import annotations

import functools
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import scipy.sparse as sp

from ..config import FEASIBILITY_TOL
from ..problem import Problem
from . import scaling as S

ROW_EQ, ROW_LE = 0, 1


@dataclass
class HostForm:
    kind: str                      # "box" | "reference"
    K: sp.csr_matrix               # m x n, canonical CSR (sorted, no duplicates), int32 indices
    c: np.ndarray                  # n
    b: np.ndarray                  # m
    lo: Optional[np.ndarray]       # n or None (= 0)
    hi: Optional[np.ndarray]       # n or None (= +inf)
    row_kind: Optional[np.ndarray]  # int8 m or None (all '=')
    b_shift: np.ndarray            # m
    lo_shift: Optional[np.ndarray]  # n or None
    ub_span: Optional[np.ndarray]  # n or None
    norm_b: float
    norm_c: float
    n_orig: int
    m_ub: int
    m_eq: int
    # reference form only
    lb_f: Optional[np.ndarray] = None
    row_sign: Optional[np.ndarray] = None
    n_ub_rows: int = 0
    # scaled box form only: K = diag(row_scale) K0 diag(col_scale)
    row_scale: Optional[np.ndarray] = None
    col_scale: Optional[np.ndarray] = None
    meta: dict = field(default_factory=dict)

    @property
    def m(self) -> int:
        return self.K.shape[0]

    @functools.cached_property
    def KT(self) -> sp.csr_matrix:
        """n x m, built on the host on first use (the engines build K^T on the device)."""
        return _transpose(self.K)

    @property
    def n(self) -> int:
        return self.K.shape[1]

    @property
    def nnz(self) -> int:
        return self.K.nnz


def _canonical(A) -> sp.csr_matrix:
    A = sp.csr_matrix(A, dtype=np.float64)
    A.sum_duplicates()
    A.sort_indices()
    if A.nnz >= 2**31 - 1:
        raise ValueError("nnz beyond int32 indexing")
    A.indices = A.indices.astype(np.int32, copy=False)
    A.indptr = A.indptr.astype(np.int32, copy=False)
    return A


def _transpose(A: sp.csr_matrix) -> sp.csr_matrix:
    return _canonical(A.T.tocsr())


def box_form(p: Problem, scaling: str = "none") -> HostForm:
    n = p.n_vars
    m_ub, m_eq = p.n_ineq, p.n_eq
    K = _canonical(sp.vstack([sp.csr_matrix(p.A_ub), sp.csr_matrix(p.A_eq)], format="csr")
                   if (m_ub + m_eq) else sp.csr_matrix((0, n)))
    b = np.concatenate([np.asarray(p.b_ub, float), np.asarray(p.b_eq, float)])
    lb = np.asarray(p.lb, float)
    ub = np.asarray(p.ub, float)
    lb_f = np.where(np.isfinite(lb), lb, 0.0)
    fin_ub = np.isfinite(ub)
    kind = np.concatenate([np.full(m_ub, ROW_LE, np.int8), np.full(m_eq, ROW_EQ, np.int8)])
    b_shift = b - K @ lb_f if lb_f.any() else b.copy()
    span = np.where(fin_ub, ub - lb_f, np.inf)
    norm_b = max(1.0, float(np.sqrt(b_shift @ b_shift + np.sum(span[fin_ub] ** 2))))
    c = np.asarray(p.c, float)
    norm_c = max(1.0, float(np.linalg.norm(c)))
    Dr = Dc = None
    if scaling != "none":
        Dr, Dc = S.scales(K, scaling)
        K = _canonical(S.scale_matrix(K, Dr, Dc))
        b, b_shift, c = b * Dr, b_shift * Dr, c * Dc
        lb, ub, lb_f, span = lb / Dc, ub / Dc, lb_f / Dc, span / Dc
    elif scaling not in S.METHODS:
        raise ValueError(f"scaling must be one of {S.METHODS}, got {scaling!r}")
    trivial_box = bool(np.all(lb == 0.0) and not fin_ub.any())
    return HostForm(
        kind="box", K=K, c=c, b=b,
        lo=None if trivial_box else lb, hi=None if trivial_box else ub,
        row_kind=kind if m_ub else None,
        b_shift=b_shift, lo_shift=None if not lb_f.any() else lb_f,
        ub_span=span if fin_ub.any() else None,
        norm_b=norm_b, norm_c=norm_c,
        n_orig=n, m_ub=m_ub, m_eq=m_eq, lb_f=lb_f, row_scale=Dr, col_scale=Dc,
    )


def reference_form(p: Problem) -> HostForm:
    """Equality standard form of ``_to_standard_form_pdhg`` (pdhg.py:62-125)."""
    n = p.n_vars
    lb = np.asarray(p.lb, float)
    ub = np.asarray(p.ub, float)
    lb_f = np.where(np.isfinite(lb), lb, 0.0)
    F = np.flatnonzero(np.isfinite(ub))
    nF = F.size
    A_ub = sp.csr_matrix(p.A_ub)
    A_eq = sp.csr_matrix(p.A_eq)
    rows_le = sp.vstack([A_ub, sp.csr_matrix((np.ones(nF), (np.arange(nF), F)), shape=(nF, n))], format="csr")
    b_le = np.concatenate([np.asarray(p.b_ub, float) - A_ub @ lb_f, (ub - lb_f)[F]])
    m_le = rows_le.shape[0]
    m_eq = A_eq.shape[0]
    n_full = n + m_le
    blocks = []
    if m_le:
        blocks.append(sp.hstack([rows_le, sp.identity(m_le, format="csr")], format="csr"))
    if m_eq:
        blocks.append(sp.hstack([A_eq, sp.csr_matrix((m_eq, m_le))], format="csr"))
    K = sp.vstack(blocks, format="csr") if blocks else sp.csr_matrix((0, n_full))
    b = np.concatenate([b_le, np.asarray(p.b_eq, float) - A_eq @ lb_f])
    sign = np.where(b < -FEASIBILITY_TOL, -1.0, 1.0)
    if (sign < 0).any():
        K = sp.diags(sign) @ K
        b = b * sign
    K = _canonical(K)
    c = np.concatenate([np.asarray(p.c, float), np.zeros(m_le)])
    return HostForm(
        kind="reference", K=K, c=c, b=b, lo=None, hi=None, row_kind=None,
        b_shift=b.copy(), lo_shift=None, ub_span=None,
        norm_b=max(1.0, float(np.linalg.norm(b))), norm_c=max(1.0, float(np.linalg.norm(c))),
        n_orig=n, m_ub=p.n_ineq, m_eq=m_eq, lb_f=lb_f, row_sign=sign, n_ub_rows=nF,
    )


def make_form(p: Problem, formulation: str, scaling: str = "none") -> HostForm:
    if formulation == "box":
        return box_form(p, scaling)
    if formulation == "reference":
        if scaling != "none":
            raise ValueError("scaling applies to formulation='box' (the reference form is the parity mode)")
        return reference_form(p)
    raise ValueError(f"formulation must be 'box' or 'reference', got {formulation!r}")


def extract(form: HostForm, p: Problem, x_dev: np.ndarray, y_dev: np.ndarray):
    """Map engine (x, y) back to the original space (pdhg.py:246-259).

    Returns ``(x, objective, y)`` with y ordered ``[A_ub rows; A_eq rows]`` in
    the sign convention of ``L = cᵀx + yᵀ(b − Ax)`` (y ≤ 0 on ``<=`` rows).
    """
    if form.kind == "box":
        x = np.asarray(x_dev[: form.n_orig], float)
        y = np.asarray(y_dev, float)
        if form.col_scale is not None:
            x = x * form.col_scale
            y = y * form.row_scale
    else:
        x = np.asarray(x_dev[: form.n_orig], float) + form.lb_f
        ys = np.asarray(y_dev, float) * form.row_sign
        y = np.concatenate([ys[: form.m_ub], ys[form.m_ub + form.n_ub_rows:]])
    x = np.clip(x, np.where(np.isfinite(p.lb), p.lb, -1e30), np.where(np.isfinite(p.ub), p.ub, 1e30))
    obj = float(np.asarray(p.c, float) @ x)
    if p.sense == "max":
        obj = -obj
    return x, obj, y
