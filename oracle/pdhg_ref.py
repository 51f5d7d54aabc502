"""FP64 CPU restatement of the reference restarted PDHG.  TEST INFRASTRUCTURE.

Follows ``solver/lp/pdhg.py`` step by step (the algorithm has to be the same
for the published iteration counts to be reproduced); the code is written
afresh.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU
baseline use it.

Pinned by: the reference's published PDHG results (status, objective,
iteration count) on afiro, blending, brandy, adlittle and the three seeded
stress LPs — ``benchmarks/results/benchmark_comparison.csv:5,9,13,17,27,31,35``
— and by the closed-form checks in ``tests/test_oracle.py`` (standard-form
sizes and feasibility-preserving lift, step-size bound ‖Σ½AT½‖₂ ≤ 0.999).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

FEASIBILITY_TOL = 1e-6       # solver/config.py:10
PDHG_TOL = 1e-6              # solver/config.py:12
MAX_PDHG_ITERS = 100_000     # solver/config.py:37
CHECK_INTERVAL = 50          # pdhg.py:207 with PDHG_RESTART_PERIOD=50 (config.py:38)


@dataclass
class RefForm:
    """The reference's equality standard form ``min cᵀz, Az = b, z ≥ 0``."""
    c: np.ndarray
    A: sp.csr_matrix
    b: np.ndarray
    lb_shift: np.ndarray     # lb with -inf replaced by 0 (pdhg.py:72-73)
    n_orig: int
    row_sign: np.ndarray     # +1 / -1 per row (pdhg.py:116-122)
    ub_cols: np.ndarray      # columns that received an upper-bound row (pdhg.py:78-82)


def to_standard_form(p) -> RefForm:
    """``_to_standard_form_pdhg`` (pdhg.py:62-125).

    rows ``[A_ub; E_F; A_eq]``, columns ``[x; slacks of the ≤ rows]``; shift by
    the finite lower bounds; negate rows whose right-hand side is below
    ``-FEASIBILITY_TOL``.
    """
    n = p.n_vars
    lbf = np.where(np.isfinite(p.lb), p.lb, 0.0)
    ub_cols = np.flatnonzero(np.isfinite(p.ub) & np.isfinite(lbf))
    k = ub_cols.size
    E = sp.csr_matrix((np.ones(k), (np.arange(k), ub_cols)), shape=(k, n))
    A_le = sp.vstack([p.A_ub, E], format="csr") if k else sp.csr_matrix(p.A_ub)
    b_le = np.concatenate([p.b_ub - p.A_ub @ lbf, (p.ub - lbf)[ub_cols]])
    m_le = A_le.shape[0]
    m_eq = p.A_eq.shape[0]
    top = sp.hstack([A_le, sp.identity(m_le, format="csr")], format="csr") if m_le \
        else sp.csr_matrix((0, n + m_le))
    bottom = sp.hstack([p.A_eq, sp.csr_matrix((m_eq, m_le))], format="csr") if m_eq \
        else sp.csr_matrix((0, n + m_le))
    A = sp.vstack([top, bottom], format="csr")
    b = np.concatenate([b_le, p.b_eq - p.A_eq @ lbf])
    sign = np.where(b < -FEASIBILITY_TOL, -1.0, 1.0)
    if (sign < 0).any():
        A = sp.csr_matrix(sp.diags(sign) @ A)
        b = b * sign
    A.sum_duplicates()
    A.sort_indices()
    c = np.concatenate([np.asarray(p.c, float), np.zeros(m_le)])
    return RefForm(c=c, A=A, b=b, lb_shift=lbf, n_orig=n, row_sign=sign, ub_cols=ub_cols)


def step_sizes(A: sp.csr_matrix):
    """``_compute_step_sizes`` (pdhg.py:131-147): τ_j = .999/max(Σ_i|A_ij|,1e-4), σ_i likewise on rows."""
    absA = abs(A)
    col = np.maximum(np.asarray(absA.sum(axis=0)).ravel(), 1e-4)
    row = np.maximum(np.asarray(absA.sum(axis=1)).ravel(), 1e-4)
    return 0.999 / col, 0.999 / row


def kkt(A, AT, b, c, norm_b, norm_c, x, y):
    """``_kkt_residuals`` (pdhg.py:184-196): (max, rp, rd, gap)."""
    rp = float(np.linalg.norm(A @ x - b)) / norm_b
    r = c - AT @ y
    rd = float(np.linalg.norm(np.minimum(r, 0.0))) / norm_c
    po = float(c @ x)
    do = float(b @ y)
    gap = abs(po - do) / (1.0 + abs(po) + abs(do))
    return max(rp, rd, gap), rp, rd, gap


@dataclass
class RefResult:
    status: str
    x: np.ndarray
    objective: float
    iterations: int
    z: np.ndarray = None         # standard-form primal at exit
    y: np.ndarray = None         # standard-form dual at exit
    restarts: int = 0
    last_err: float = float("nan")


def solve(p, max_iters: int = MAX_PDHG_ITERS, tol: float = PDHG_TOL, trace=None) -> RefResult:
    """``solve_lp_pdhg`` (pdhg.py:152-259), FP64.

    ``trace`` (a list, optional) receives ``(iteration, err, restarted)`` at
    every check, for trajectory comparisons.
    """
    sf = to_standard_form(p)
    A, b, c = sf.A, sf.b, sf.c
    m, n = A.shape
    if m == 0 or n == 0:                                  # pdhg.py:171-172
        return RefResult("optimal", np.zeros(p.n_vars), 0.0, 0)
    tau, sigma = step_sizes(A)
    AT = A.T.tocsr()
    norm_b = max(1.0, float(np.linalg.norm(b)))           # pdhg.py:181-182
    norm_c = max(1.0, float(np.linalg.norm(c)))
    x = np.zeros(n)
    y = np.zeros(m)
    xs = np.zeros(n)
    ys = np.zeros(m)
    win = 0
    ref_err = kkt(A, AT, b, c, norm_b, norm_c, x, y)[0]   # pdhg.py:203
    status = "iteration_limit"
    it = 0
    restarts = 0
    last = float("nan")
    for it in range(1, max_iters + 1):                    # pdhg.py:209-244
        xn = np.maximum(x - tau * (c - AT @ y), 0.0)
        xb = 2.0 * xn - x
        y = y + sigma * (b - A @ xb)
        x = xn
        xs += x
        ys += y
        win += 1
        if win % CHECK_INTERVAL == 0:
            xa = xs / win
            ya = ys / win
            err = kkt(A, AT, b, c, norm_b, norm_c, xa, ya)[0]
            last = err
            if err < tol:
                x, y = xa, ya
                status = "optimal"
                if trace is not None:
                    trace.append((it, err, False))
                break
            restarted = err < 0.5 * ref_err
            if restarted:
                x, y = xa.copy(), ya.copy()
                xs[:] = 0.0
                ys[:] = 0.0
                win = 0
                ref_err = err
                restarts += 1
            if trace is not None:
                trace.append((it, err, restarted))
    xo = np.clip(x[:sf.n_orig] + sf.lb_shift,                # pdhg.py:246-259
                 np.where(np.isfinite(p.lb), p.lb, -1e30),
                 np.where(np.isfinite(p.ub), p.ub, 1e30))
    obj = float(p.c @ xo)
    if p.sense == "max":
        obj = -obj
    return RefResult(status, xo, obj, it, z=x, y=y, restarts=restarts, last_err=last)
