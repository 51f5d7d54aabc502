"""Free-format MPS reader producing a :class:`Problem`.

The resulting ``Problem`` must be the one the reference reader builds
(``solver/io/mps_reader.py:42-331``), because the PDHG trajectory — and so
the published iteration counts this engine is checked against — depends on
the row order and on which rows are negated.  The rules kept from the
reference are:

* the first ``N`` row is the objective; later ``N`` rows are dropped (:141-147, :290-316);
* ``L`` rows go to ``A_ub`` in file order; ``G`` rows go to ``A_ub`` negated (:273-280);
* a ``RANGES`` entry on an ``L`` row appends the negated lower row right after it;
  on a ``G`` row it appends the (un-negated) upper row; an ``E`` row with a range becomes
  an upper row followed by a negated lower row at ``rhs`` (:262-287);
* repeated (row, column) entries are summed (:230-238);
* bounds ``UP LO FX FR MI PL BV`` as in :185-209, defaults ``lb=0, ub=+inf``;
  ``MI`` leaves the upper bound as it was; unknown bound types are ignored;
* ``INTORG``/``INTEND`` markers and ``BV`` set ``integer_mask``;
* ``$`` starts a comment; lines starting with ``*`` are comments.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import scipy.sparse as sp

from ..problem import Problem


def _fields(line: str) -> list:
    return line.split("$", 1)[0].split()


def read_mps(path) -> Problem:
    path = Path(path)
    if not path.exists():
        raise FileNotFoundError(f"MPS file not found: {path}")

    name = ""
    kinds: dict = {}             # row name -> 'N' | 'L' | 'G' | 'E'
    rows_in_order: list = []     # constraint rows (and later N rows), file order
    objective_row = None
    col_of: dict = {}
    n_cols = 0
    coef: dict = {}              # row name -> {col: value}
    obj: dict = {}
    rhs: dict = {}
    rng: dict = {}
    lo_set: dict = {}
    hi_set: dict = {}
    integer_cols: set = set()
    in_int = False
    section = ""

    def column(vname: str) -> int:
        nonlocal n_cols
        j = col_of.get(vname)
        if j is None:
            j = col_of[vname] = n_cols
            n_cols += 1
        return j

    for raw in path.read_text(encoding="utf-8", errors="replace").splitlines():
        s = raw.strip()
        if not s or s.startswith("*"):
            continue
        if raw[0] not in " \t":
            f = _fields(raw)
            if not f:
                continue
            section = f[0].upper()
            if section == "NAME":
                name = f[1] if len(f) > 1 else ""
            elif section == "ENDATA":
                break
            continue
        f = _fields(raw)
        if not f:
            continue
        if section == "ROWS":
            kind, rname = f[0].upper(), f[1]
            kinds[rname] = kind
            if kind == "N" and objective_row is None:
                objective_row = rname
            else:
                rows_in_order.append(rname)
                coef[rname] = {}
        elif section == "COLUMNS":
            if len(f) >= 2 and f[1].strip("'").upper() == "MARKER":
                if len(f) >= 3:
                    tag = f[2].strip("'").upper()
                    if tag == "INTORG":
                        in_int = True
                    elif tag == "INTEND":
                        in_int = False
                continue
            j = column(f[0])
            if in_int:
                integer_cols.add(j)
            for rname, val in zip(f[1::2], f[2::2]):
                v = float(val)
                if rname == objective_row:
                    obj[j] = obj.get(j, 0.0) + v
                elif rname in coef:
                    coef[rname][j] = coef[rname].get(j, 0.0) + v
        elif section == "RHS":
            for rname, val in zip(f[1::2], f[2::2]):
                rhs[rname] = float(val)
        elif section == "RANGES":
            for rname, val in zip(f[1::2], f[2::2]):
                rng[rname] = float(val)
        elif section == "BOUNDS":
            if len(f) < 3:
                continue
            btype = f[0].upper()
            j = column(f[2])
            val = float(f[3]) if len(f) > 3 else 0.0
            if btype == "UP":
                hi_set[j] = val
            elif btype == "LO":
                lo_set[j] = val
            elif btype == "FX":
                lo_set[j] = hi_set[j] = val
            elif btype == "FR":
                lo_set[j], hi_set[j] = -np.inf, np.inf
            elif btype == "MI":
                lo_set[j] = -np.inf
            elif btype == "BV":
                lo_set[j], hi_set[j] = 0.0, 1.0
                integer_cols.add(j)

    if n_cols == 0:
        raise ValueError("No variables found in MPS file.")

    ub_rows, eq_rows = [], []     # (coef dict, rhs, sign)
    for rname in rows_in_order:
        kind = kinds[rname]
        cf = coef[rname]
        r = rhs.get(rname, 0.0)
        R = rng.get(rname)
        if kind == "L":
            ub_rows.append((cf, r, 1.0))
            if R is not None:
                ub_rows.append((cf, -(r - abs(R)), -1.0))
        elif kind == "G":
            ub_rows.append((cf, -r, -1.0))
            if R is not None:
                ub_rows.append((cf, r + abs(R), 1.0))
        elif kind == "E":
            if R is not None:
                ub_rows.append((cf, r + abs(R), 1.0))
                ub_rows.append((cf, -r, -1.0))
            else:
                eq_rows.append((cf, r, 1.0))

    def assemble(rows):
        m = len(rows)
        if m == 0:
            return sp.csr_matrix((0, n_cols), dtype=float), np.zeros(0)
        ri, ci, vv = [], [], []
        b = np.empty(m)
        for i, (cf, r, sgn) in enumerate(rows):
            b[i] = r
            for j, v in cf.items():
                ri.append(i)
                ci.append(j)
                vv.append(sgn * v)
        A = sp.csr_matrix((np.asarray(vv, float), (np.asarray(ri, np.int64), np.asarray(ci, np.int64))),
                          shape=(m, n_cols))
        A.sum_duplicates()
        A.sort_indices()
        return A, b

    A_ub, b_ub = assemble(ub_rows)
    A_eq, b_eq = assemble(eq_rows)
    c = np.zeros(n_cols)
    for j, v in obj.items():
        c[j] = v
    lb = np.zeros(n_cols)
    ub = np.full(n_cols, np.inf)
    for j, v in lo_set.items():
        lb[j] = v
    for j, v in hi_set.items():
        ub[j] = v
    imask = np.zeros(n_cols, dtype=bool)
    imask[list(integer_cols)] = True
    p = Problem(c=c, A_ub=A_ub, b_ub=b_ub, A_eq=A_eq, b_eq=b_eq, lb=lb, ub=ub,
                integer_mask=imask, sense="min", name=name)
    p.validate()
    return p
