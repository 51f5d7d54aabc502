"""Seeded LP families of the measured workloads.

Every generator is vectorised NumPy/SciPy and deterministic in its seed
(``np.random.SeedSequence([family, seed])``).

* :func:`random_sparse_lp` — min cᵀx, Ax = b, 0 ≤ x ≤ u; A with
  ``round(density·m·n)`` N(0,1) entries at distinct random positions,
  ``u ~ U[1,10]``, feasible by construction (``x0 ~ U[0,1]·u``, ``b = A x0``)
  and bounded (``c = Aᵀy0 + s0``, ``y0 ~ N(0,1)``, ``s0 = |N(0,1)|``).
  m=200, n=500, density 2% is the CPU-runnable single-LP case.
* :func:`transportation_lp` — S sources × D sinks, supply ≤ rows and demand
  = rows in the unbalanced layout of ``demos/transportation_lp.py:136-146``,
  flat arc index ``i·D + j`` (``:105,112-115``); supply ~ U[50,150], demand
  scaled to 95% of total supply, cost ~ U[1,100].  S=1000, D=2000 gives
  n = 2M, m = 3,000, nnz = 4M.
* :func:`uneven_batch` — B random sparse LPs with ``m`` log-uniform in
  [m_lo, m_hi], ``n = 2.5 m``, density 1%.
* :func:`dispatch_family` — economic-dispatch LPs (generators × periods,
  demand ≥ rows, ramp ≤ rows, 0 ≤ p ≤ Pmax) sharing one matrix; per instance
  demand and cost perturbed ±20%.
* :func:`set_cover_nodes` — LP relaxations of a set-covering MILP
  (``-Ax ≤ -1``, 0 ≤ x ≤ 1, as ``benchmarks/miplib_benchmark.py:133-157``)
  and branch-and-bound nodes that differ only in variable fixings, built as
  ``solver/milp/branch_and_bound.py:95-113`` builds node problems.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .problem import Problem


def _rng(family: int, seed: int) -> np.random.Generator:
    return np.random.default_rng(np.random.SeedSequence([family, seed]))


def _empty(n: int) -> sp.csr_matrix:
    return sp.csr_matrix((0, n), dtype=float)


def random_sparse_lp(m: int = 200, n: int = 500, density: float = 0.02, seed: int = 0,
                     family: int = 0) -> Problem:
    rng = _rng(family, seed)
    k = max(1, int(round(density * m * n)))
    pos = np.sort(rng.choice(m * n, size=k, replace=False))
    vals = rng.standard_normal(k)
    A = sp.csr_matrix((vals, (pos // n, pos % n)), shape=(m, n))
    u = rng.uniform(1.0, 10.0, n)
    x0 = rng.uniform(0.0, 1.0, n) * u
    b = A @ x0
    y0 = rng.standard_normal(m)
    s0 = np.abs(rng.standard_normal(n))
    c = A.T @ y0 + s0
    return Problem(c=c, A_ub=_empty(n), b_ub=np.zeros(0), A_eq=A, b_eq=b, lb=np.zeros(n), ub=u,
                   integer_mask=np.zeros(n, dtype=bool), name=f"rand_m{m}_n{n}_s{seed}")


def transportation_lp(S: int = 1000, D: int = 2000, seed: int = 0) -> Problem:
    rng = _rng(1, seed)
    supply = rng.uniform(50.0, 150.0, S)
    d_raw = rng.uniform(50.0, 150.0, D)
    demand = d_raw * (0.95 * supply.sum() / d_raw.sum())
    cost = rng.uniform(1.0, 100.0, S * D)
    n = S * D
    ones = np.ones(n)
    A_ub = sp.csr_matrix((ones, np.arange(n, dtype=np.int64), np.arange(S + 1, dtype=np.int64) * D), shape=(S, n))
    cols = (np.arange(D, dtype=np.int64)[:, None] + np.arange(S, dtype=np.int64)[None, :] * D).ravel()
    A_eq = sp.csr_matrix((ones, cols, np.arange(D + 1, dtype=np.int64) * S), shape=(D, n))
    return Problem(c=cost, A_ub=A_ub, b_ub=supply, A_eq=A_eq, b_eq=demand, lb=np.zeros(n),
                   ub=np.full(n, np.inf), integer_mask=np.zeros(n, dtype=bool),
                   name=f"transport_{S}x{D}_s{seed}")


def uneven_sizes(B: int, m_lo: int = 50, m_hi: int = 2000, seed: int = 0) -> np.ndarray:
    rng = _rng(2, 10**9 + seed)
    return np.rint(np.exp(rng.uniform(np.log(m_lo), np.log(m_hi), B))).astype(int)


def uneven_batch(B: int = 4096, m_lo: int = 50, m_hi: int = 2000, density: float = 0.01,
                 seed: int = 0) -> list:
    ms = uneven_sizes(B, m_lo, m_hi, seed)
    return [random_sparse_lp(int(m), int(round(2.5 * m)), density, seed=seed * 1_000_003 + k, family=2)
            for k, m in enumerate(ms)]
