"""Seeded stress LPs whose PDHG results the reference publishes.

Rebuilds the three instances of ``benchmarks/run_benchmarks.py:321-375``
(random_n50_m30, sparse_n500_m250, badly_scaled_1e8).  They share one
``default_rng(42)`` stream, so the draws must happen in the reference's order:
dense uniform A, x0, b noise, c; then ``sp.random(..., random_state=42)`` with
values drawn from the shared stream, x_true, b noise, c.  Expected outputs are
in ``pdhg_published.csv`` next to this file.
"""

from __future__ import annotations

import os

import numpy as np
import scipy.sparse as sp

from sovereign_solver_amd.io.mps import read_mps
from sovereign_solver_amd.problem import Problem

HERE = os.path.dirname(os.path.abspath(__file__))


def _lp(c, A_ub, b_ub, lb, ub, name):
    n = len(c)
    return Problem(c=np.asarray(c, float), A_ub=sp.csr_matrix(A_ub), b_ub=np.asarray(b_ub, float),
                   A_eq=sp.csr_matrix((0, n)), b_eq=np.zeros(0), lb=np.asarray(lb, float),
                   ub=np.asarray(ub, float), integer_mask=np.zeros(n, dtype=bool), name=name)


def stress_instances():
    rng = np.random.default_rng(42)
    n, m = 50, 30
    A = rng.uniform(0.1, 2.0, (m, n))
    x0 = rng.uniform(0.5, 2.0, n)
    b = A @ x0 + rng.uniform(0.1, 1.0, m)
    c = rng.standard_normal(n)
    med = _lp(c, np.vstack([A, np.eye(n)]), np.concatenate([b, np.full(n, 10.0)]),
              np.zeros(n), np.full(n, np.inf), "random_n50_m30")

    n2, m2 = 500, 250
    A2 = sp.random(m2, n2, density=0.03, random_state=42, data_rvs=lambda s: rng.uniform(0.5, 2.0, s))
    xt = rng.uniform(0.5, 3.0, n2)
    b2 = A2 @ xt + rng.uniform(0.1, 1.0, m2)
    c2 = rng.uniform(-5.0, 5.0, n2)
    big = _lp(c2, sp.vstack([A2, sp.eye(n2, format="csr")]), np.concatenate([b2, np.full(n2, 10.0)]),
              np.zeros(n2), np.full(n2, np.inf), "sparse_n500_m250")

    bad = _lp(np.array([1e-4, 1e2, 1e-2, 1e4]),
              np.array([[1e4, 1e-2, 1.0, 1e-3], [1e-1, 1e3, 1e-4, 1e2], [1.0, 1.0, 1e4, 1e-1]]),
              np.array([1e5, 1e4, 1e4]), np.zeros(4), np.full(4, 100.0), "badly_scaled_1e8")
    return {p.name: p for p in (med, big, bad)}


def netlib_instance(name: str) -> Problem:
    return read_mps(os.path.join(HERE, "instances", f"{name}.mps"))


def load_published():
    import csv
    out = {}
    with open(os.path.join(HERE, "pdhg_published.csv")) as f:
        for row in csv.DictReader(f):
            out[(row["problem"], row["kind"])] = row
    return out


def golden_problem(name: str) -> Problem:
    if name in ("afiro", "blending", "brandy", "adlittle"):
        return netlib_instance(name)
    return stress_instances()[name]
