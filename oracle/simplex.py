"""FP64 bounded-variable revised primal simplex.  TEST INFRASTRUCTURE.

The objective oracle of the PDHG results: an exact (vertex) method, written
here from the textbook algorithm (two phases, bounded variables, Dantzig
pricing with Bland's rule after a run of degenerate pivots, explicit basis
inverse with eta updates and periodic refactorisation).  It does not follow
the reference's own simplex (``solver/lp/simplex_revised.py``, Big-M with
dense artificial rows); only its role is the same.

Pinned in ``tests/test_simplex_oracle.py`` against
``scipy.optimize.linprog(method="highs-ds")`` and against brute-force vertex
enumeration on tiny LPs.  Dense linear algebra: meant for LPs with up to a
few hundred rows.

Form: min cᵀx, A_ub x ≤ b_ub, A_eq x = b_eq, lb ≤ x ≤ ub, written as
``[A_ub I; A_eq 0] [x; s] = b`` with slacks s ≥ 0 and one artificial per row
for phase 1.  Duals are returned in the convention L = cᵀx + yᵀ(b − Ax)
(y ≤ 0 on ``<=`` rows), rows ordered ``[A_ub; A_eq]``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

TOL_P = 1e-9     # primal feasibility
TOL_D = 1e-9     # reduced-cost optimality
TOL_PIV = 1e-10  # smallest usable pivot


@dataclass
class SimplexResult:
    status: str          # optimal | infeasible | unbounded | iteration_limit
    x: np.ndarray
    objective: float
    y: np.ndarray
    iterations: int


class _Simplex:
    def __init__(self, A: np.ndarray, b: np.ndarray, lo: np.ndarray, hi: np.ndarray):
        self.A = A
        self.b = b
        self.lo = lo.copy()
        self.hi = hi.copy()
        self.m, self.n = A.shape

    def _refactor(self):
        self.Binv = np.linalg.inv(self.A[:, self.basis])
        self.since = 0

    def _xb(self, x):
        nb = np.ones(self.n, bool)
        nb[self.basis] = False
        rhs = self.b - self.A[:, nb] @ x[nb]
        return self.Binv @ rhs

    def run(self, c: np.ndarray, x: np.ndarray, basis: np.ndarray, max_iter: int):
        """Primal simplex from a feasible basis; x holds nonbasic values at bounds."""
        self.basis = basis.copy()
        self._refactor()
        m = self.m
        degenerate = 0
        it = 0
        while it < max_iter:
            it += 1
            if self.since >= 64:
                self._refactor()
            x[self.basis] = self._xb(x)
            y = self.Binv.T @ c[self.basis]
            d = c - self.A.T @ y
            d[self.basis] = 0.0
            at_lo = np.isclose(x, self.lo, rtol=0, atol=TOL_P) & np.isfinite(self.lo)
            at_hi = np.isclose(x, self.hi, rtol=0, atol=TOL_P) & np.isfinite(self.hi)
            fixed = at_lo & at_hi
            nonbasic = np.ones(self.n, bool)
            nonbasic[self.basis] = False
            can_up = nonbasic & ~fixed & ~at_hi & (d < -TOL_D)
            can_dn = nonbasic & ~fixed & ~at_lo & (d > TOL_D)
            cand = np.flatnonzero(can_up | can_dn)
            if cand.size == 0:
                return "optimal", x, y, it
            if degenerate >= 50:
                j = int(cand[0])                               # Bland
            else:
                j = int(cand[np.argmax(np.abs(d[cand]))])      # Dantzig
            delta = 1.0 if d[j] < 0 else -1.0
            w = self.Binv @ self.A[:, j]
            xb = x[self.basis]
            lb, ub = self.lo[self.basis], self.hi[self.basis]
            t_best = self.hi[j] - self.lo[j]                    # bound flip
            r = -1
            dw = delta * w
            for i in range(m):
                if dw[i] > TOL_PIV and np.isfinite(lb[i]):
                    t = (xb[i] - lb[i]) / dw[i]
                elif dw[i] < -TOL_PIV and np.isfinite(ub[i]):
                    t = (ub[i] - xb[i]) / (-dw[i])
                else:
                    continue
                t = max(t, 0.0)
                if t < t_best - 1e-12 or (r >= 0 and abs(t - t_best) <= 1e-12 and abs(w[i]) > abs(w[r])):
                    t_best, r = t, i
            if not np.isfinite(t_best):
                return "unbounded", x, y, it
            degenerate = degenerate + 1 if t_best <= 1e-12 else 0
            x[j] += delta * t_best
            x[self.basis] = xb - dw * t_best
            if r < 0:
                continue                                        # entering variable flipped bound
            leave = self.basis[r]
            x[leave] = lb[r] if dw[r] > 0 else ub[r]
            # eta update of the basis inverse
            piv = w[r]
            row = self.Binv[r] / piv
            self.Binv -= np.outer(w, row)
            self.Binv[r] = row
            self.basis[r] = j
            self.since += 1
        return "iteration_limit", x, None, it


def solve(p, max_iter: int = 50_000) -> SimplexResult:
    n = p.n_vars
    mu, me = p.n_ineq, p.n_eq
    m = mu + me
    A_ub = p.A_ub.toarray() if sp.issparse(p.A_ub) else np.asarray(p.A_ub, float)
    A_eq = p.A_eq.toarray() if sp.issparse(p.A_eq) else np.asarray(p.A_eq, float)
    A = np.zeros((m, n + mu))
    A[:mu, :n] = A_ub.reshape(mu, n)
    A[mu:, :n] = A_eq.reshape(me, n)
    A[:mu, n:] = np.eye(mu)
    b = np.concatenate([np.asarray(p.b_ub, float), np.asarray(p.b_eq, float)])
    lo = np.concatenate([np.asarray(p.lb, float), np.zeros(mu)])
    hi = np.concatenate([np.asarray(p.ub, float), np.full(mu, np.inf)])
    c = np.concatenate([np.asarray(p.c, float), np.zeros(mu)])
    nt = n + mu
    if m == 0:
        x = np.where(c > 0, lo, np.where(c < 0, hi, np.where(np.isfinite(lo), lo, np.where(np.isfinite(hi), hi, 0.0))))
        if not np.all(np.isfinite(x)):
            return SimplexResult("unbounded", x[:n], -np.inf, np.zeros(0), 0)
        obj = float(c @ x)
        return SimplexResult("optimal", x[:n], -obj if p.sense == "max" else obj, np.zeros(0), 0)
    # starting nonbasic values: a finite bound (lower first), free variables at 0
    x0 = np.where(np.isfinite(lo), lo, np.where(np.isfinite(hi), hi, 0.0))
    r = b - A @ x0
    sgn = np.where(r >= 0, 1.0, -1.0)
    Aa = np.hstack([A, np.diag(sgn)])
    lo_a = np.concatenate([lo, np.zeros(m)])
    hi_a = np.concatenate([hi, np.full(m, np.inf)])
    x = np.concatenate([x0, np.abs(r)])
    basis = np.arange(nt, nt + m)
    S = _Simplex(Aa, b, lo_a, hi_a)
    c1 = np.concatenate([np.zeros(nt), np.ones(m)])
    st, x, _, it1 = S.run(c1, x, basis, max_iter)
    if st != "optimal":
        return SimplexResult(st, x[:n], np.nan, np.zeros(m), it1)
    if np.sum(x[nt:]) > 1e-7 * max(1.0, np.abs(b).max()):
        return SimplexResult("infeasible", x[:n], np.nan, np.zeros(m), it1)
    # phase 2: artificials pinned at zero
    S.hi[nt:] = 0.0
    x[nt:] = 0.0
    c2 = np.concatenate([c, np.zeros(m)])
    st, x, y, it2 = S.run(c2, x, S.basis, max_iter)
    obj = float(c @ x[:nt])
    if p.sense == "max":
        obj = -obj
    return SimplexResult(st, x[:n].copy(), obj, y if y is not None else np.zeros(m), it1 + it2)
