"""LP input type, interface-compatible with the reference's ``Problem``.

Mirrors ``solver/problem.py:19-67`` (fields, order and defaults),
``:71-99`` (derived sizes) and ``:103-142`` (``validate``):

    min/max  cᵀx            A_ub x ≤ b_ub,  A_eq x = b_eq,  lb ≤ x ≤ ub

``sense='max'`` follows the reference convention (``problem.py:50-52``): ``c``
is already negated by whoever built the problem and the solver flips the sign
of the reported objective.  ``integer_mask`` and ``P_qp`` are carried for API
compatibility; the PDHG path ignores them, as the reference's does.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np
import scipy.sparse as sp


@dataclass
class Problem:
    c: np.ndarray
    A_ub: sp.csr_matrix
    b_ub: np.ndarray
    A_eq: sp.csr_matrix
    b_eq: np.ndarray
    lb: np.ndarray
    ub: np.ndarray
    integer_mask: np.ndarray
    P_qp: Optional[sp.csr_matrix] = None
    sense: str = "min"
    name: str = ""

    # sizes (problem.py:71-89)
    @property
    def n_vars(self) -> int:
        return len(self.c)

    @property
    def n_ineq(self) -> int:
        return self.A_ub.shape[0]

    @property
    def n_eq(self) -> int:
        return self.A_eq.shape[0]

    @property
    def n_constraints(self) -> int:
        return self.n_ineq + self.n_eq

    @property
    def is_milp(self) -> bool:
        return bool(np.any(self.integer_mask))

    @property
    def is_qp(self) -> bool:
        return self.P_qp is not None

    def validate(self) -> None:
        """Raise ValueError listing every inconsistent dimension (problem.py:103-142)."""
        n = self.n_vars
        bad = []

        def need(ok: bool, msg: str) -> None:
            if not ok:
                bad.append(msg)

        need(self.A_ub.shape == (self.n_ineq, n), f"A_ub shape {self.A_ub.shape} inconsistent with n_vars={n}")
        need(np.shape(self.b_ub) == (self.n_ineq,), f"b_ub shape {np.shape(self.b_ub)} != ({self.n_ineq},)")
        need(self.A_eq.shape == (self.n_eq, n), f"A_eq shape {self.A_eq.shape} inconsistent with n_vars={n}")
        need(np.shape(self.b_eq) == (self.n_eq,), f"b_eq shape {np.shape(self.b_eq)} != ({self.n_eq},)")
        need(np.shape(self.lb) == (n,), f"lb shape {np.shape(self.lb)} != ({n},)")
        need(np.shape(self.ub) == (n,), f"ub shape {np.shape(self.ub)} != ({n},)")
        need(np.shape(self.integer_mask) == (n,), f"integer_mask shape {np.shape(self.integer_mask)} != ({n},)")
        need(self.sense in ("min", "max"), f"sense must be 'min' or 'max', got '{self.sense}'")
        if self.P_qp is not None:
            need(self.P_qp.shape == (n, n), f"P_qp shape {self.P_qp.shape} inconsistent with n_vars={n}")
        if np.shape(self.lb) == np.shape(self.ub):
            crossed = np.flatnonzero(np.asarray(self.lb) > np.asarray(self.ub) + 1e-12)
            need(crossed.size == 0, f"lb > ub for variables: {crossed[:5].tolist()} ...")
        if bad:
            raise ValueError("Problem validation failed:\n  " + "\n  ".join(bad))

    @classmethod
    def empty(cls, n: int) -> "Problem":
        """An n-variable problem with no rows, x ≥ 0 (problem.py:146-158)."""
        return cls(
            c=np.zeros(n),
            A_ub=sp.csr_matrix((0, n), dtype=float),
            b_ub=np.zeros(0),
            A_eq=sp.csr_matrix((0, n), dtype=float),
            b_eq=np.zeros(0),
            lb=np.zeros(n),
            ub=np.full(n, np.inf),
            integer_mask=np.zeros(n, dtype=bool),
        )

    def eval_obj(self, x: np.ndarray) -> float:
        """Objective at x in the original sense (problem.py:160-169)."""
        v = float(self.c @ x)
        if self.P_qp is not None:
            v += 0.5 * float(x @ self.P_qp.dot(x))
        return -v if self.sense == "max" else v

    def __repr__(self) -> str:
        kind = "MILP" if self.is_milp else ("QP" if self.is_qp else "LP")
        return (f"Problem(name={self.name!r}, type={kind}, n_vars={self.n_vars}, "
                f"n_ineq={self.n_ineq}, n_eq={self.n_eq})")
