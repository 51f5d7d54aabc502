"""Diagonal scaling (preconditioning) of an LP's constraint matrix.

The scaled LP has ``K_s = D_r K D_c``, ``b_s = D_r b``, ``c_s = D_c c`` and
bounds ``l / D_c``, ``u / D_c``; its solution maps back as ``x = D_c x_s`` and
``y = D_r y_s`` (the objective is unchanged).  The engine runs PDHG on the
scaled LP and evaluates the KKT yardstick on the original one
(``PdlpDev.row_scale`` / ``col_scale``), so the tolerance means the same with
and without scaling.

Methods (the reference's ``solver/utils/scaling.py``):

``ruiz``
    Ruiz equilibration (``RuizScaler``, scaling.py:42-192): each pass divides
    every row and column by the square root of its largest absolute entry,
    both computed on the same matrix; empty lines keep scale 1; at most
    ``max_iters`` = 10 passes, stopping once every line's largest entry is
    within ``tol`` = 1e-2 of 1.
``geometric``
    Geometric-mean scaling (``GeometricMeanScaler``, scaling.py:194-290):
    each of ``max_iters`` = 5 passes divides every row by the square root of
    the geometric mean of its absolute entries (entries above 1e-15), then
    every column of the row-scaled matrix the same way.
``pock_chambolle``
    One pass dividing row i by ``sqrt(sum_j |K_ij|)`` and column j by
    ``sqrt(sum_i |K_ij|)`` (Pock and Chambolle 2011, alpha = 1; the last step
    of PDLP's default scaling).  Not in the reference.
``ruiz_pc``
    ``ruiz`` followed by ``pock_chambolle`` (PDLP's default order).

The reference's ``scale_problem`` (scaling.py:303-331) scales only ``A_ub``
while its column scales also change the ``A_eq`` columns; here every method
scales the stacked ``[A_ub; A_eq]``, which keeps the scaled LP equivalent.
"""

from __future__ import annotations

from typing import Tuple

import numpy as np
import scipy.sparse as sp

METHODS = ("none", "ruiz", "geometric", "pock_chambolle", "ruiz_pc")


def _row_max(A: sp.csr_matrix) -> np.ndarray:
    """Largest |entry| of each row (0 for an empty row)."""
    out = np.zeros(A.shape[0])
    lens = np.diff(A.indptr)
    nz = lens > 0
    if nz.any():
        out[nz] = np.maximum.reduceat(np.abs(A.data), A.indptr[:-1][nz])
    return out


def _col_max(A: sp.csr_matrix) -> np.ndarray:
    out = np.zeros(A.shape[1])
    if A.nnz:
        np.maximum.at(out, A.indices, np.abs(A.data))
    return out


def _scaled(A: sp.csr_matrix, dr: np.ndarray, dc: np.ndarray) -> sp.csr_matrix:
    """diag(dr) A diag(dc), same sparsity pattern and entry order."""
    B = A.copy()
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    B.data = A.data * dr[rows] * dc[A.indices]
    return B


def ruiz(A: sp.csr_matrix, max_iters: int = 10, tol: float = 1e-2) -> Tuple[np.ndarray, np.ndarray]:
    """(D_r, D_c) of Ruiz equilibration in the infinity norm."""
    m, n = A.shape
    Dr, Dc = np.ones(m), np.ones(n)
    W = sp.csr_matrix(A, dtype=np.float64)
    empty_r = np.diff(W.indptr) == 0
    empty_c = np.bincount(W.indices, minlength=n) == 0
    for _ in range(max_iters):
        rn, cn = _row_max(W), _col_max(W)
        rn[empty_r] = 1.0                             # empty lines count as norm 1
        cn[empty_c] = 1.0
        dr = np.where(rn > 1e-15, 1.0 / np.sqrt(rn), 1.0)
        dc = np.where(cn > 1e-15, 1.0 / np.sqrt(cn), 1.0)
        W = _scaled(W, dr, dc)
        Dr *= dr
        Dc *= dc
        dev = max(float(np.max(np.abs(rn * dr * dr - 1.0))) if m else 0.0,
                  float(np.max(np.abs(cn * dc * dc - 1.0))) if n else 0.0)
        if dev < tol:
            break
    return Dr, Dc


def _geo_rows(A: sp.csr_matrix) -> np.ndarray:
    """1 / sqrt(geometric mean of |entries| > 1e-15) per row; 1 for rows without any."""
    a = np.abs(A.data)
    keep = a > 1e-15
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    cnt = np.bincount(rows[keep], minlength=A.shape[0])
    s = np.bincount(rows[keep], weights=np.log(a[keep]), minlength=A.shape[0])
    d = np.ones(A.shape[0])
    has = cnt > 0
    gm = np.exp(s[has] / cnt[has])
    d[has] = np.where(gm > 1e-15, 1.0 / np.sqrt(gm), 1.0)
    return d


def geometric(A: sp.csr_matrix, max_iters: int = 5) -> Tuple[np.ndarray, np.ndarray]:
    """(D_r, D_c) of geometric-mean scaling, rows then columns in each pass."""
    m, n = A.shape
    Dr, Dc = np.ones(m), np.ones(n)
    W = sp.csr_matrix(A, dtype=np.float64)
    for _ in range(max_iters):
        dr = _geo_rows(W)
        W = _scaled(W, dr, np.ones(n))
        Dr *= dr
        dc = _geo_rows(W.T.tocsr())
        W = _scaled(W, np.ones(m), dc)
        Dc *= dc
    return Dr, Dc


def pock_chambolle(A: sp.csr_matrix) -> Tuple[np.ndarray, np.ndarray]:
    """(D_r, D_c) = 1 / sqrt of the row and column absolute sums (1 for empty lines)."""
    a = abs(sp.csr_matrix(A, dtype=np.float64))
    r = np.asarray(a.sum(axis=1)).ravel()
    c = np.asarray(a.sum(axis=0)).ravel()
    return 1.0 / np.sqrt(np.where(r > 0.0, r, 1.0)), 1.0 / np.sqrt(np.where(c > 0.0, c, 1.0))


def scales(A: sp.csr_matrix, method: str) -> Tuple[np.ndarray, np.ndarray]:
    """(D_r, D_c) of ``method`` (one of :data:`METHODS`) for the matrix A."""
    m, n = A.shape
    if method == "none" or m == 0 or n == 0:
        return np.ones(m), np.ones(n)
    if method == "ruiz":
        return ruiz(A)
    if method == "geometric":
        return geometric(A)
    if method == "pock_chambolle":
        return pock_chambolle(A)
    if method == "ruiz_pc":
        Dr, Dc = ruiz(A)
        pr, pc = pock_chambolle(_scaled(sp.csr_matrix(A, dtype=np.float64), Dr, Dc))
        return Dr * pr, Dc * pc
    raise ValueError(f"scaling must be one of {METHODS}, got {method!r}")


def scale_matrix(A: sp.csr_matrix, Dr: np.ndarray, Dc: np.ndarray) -> sp.csr_matrix:
    """diag(D_r) A diag(D_c) in canonical CSR."""
    return _scaled(sp.csr_matrix(A, dtype=np.float64), np.asarray(Dr, float), np.asarray(Dc, float))
