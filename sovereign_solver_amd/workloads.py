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


# demand profile shape of demos/power_economic_dispatch.py:78-83 (24 hourly values)
_DEMAND_SHAPE = np.array([300, 280, 270, 260, 280, 350, 450, 600, 700, 750, 780, 800,
                          790, 800, 780, 750, 720, 700, 680, 700, 720, 680, 600, 450], float)


class DispatchFamily:
    """Parametric economic-dispatch LPs sharing one constraint matrix.

    G generators × T periods, variables p[t*G + g] in [0, Pmax_g]; rows
    (all ``<=``): T demand rows ``-Σ_g p_gt ≤ -D_t`` and, per generator and
    period step, ramp-up ``p_gt − p_g,t−1 ≤ R_g`` and ramp-down
    ``p_g,t−1 − p_gt ≤ R_g`` with R = 0.25·Pmax.  Base data (seed 0):
    Pmax ~ U[50,400], cost ~ U[20,60], demand shaped like the reference demo's
    profile scaled to a peak of 0.7·ΣPmax.  Instance k multiplies each D_t and
    each cost_g by 1 + U[−0.2, 0.2] (seed [3, k]).  p = Pmax is feasible for
    every instance, so all are feasible and bounded.
    """

    def __init__(self, G: int = 200, T: int = 24, base_seed: int = 0):
        rng = _rng(3, 10**9 + base_seed)
        self.G, self.T = G, T
        self.pmax = rng.uniform(50.0, 400.0, G)
        self.ramp = 0.25 * self.pmax
        self.cost = rng.uniform(20.0, 60.0, G)
        shape = np.interp(np.linspace(0, 23, T), np.arange(24), _DEMAND_SHAPE)
        self.demand = shape / shape.max() * 0.7 * self.pmax.sum()
        n = G * T
        rows, cols, vals = [], [], []
        for t in range(T):                                   # demand rows
            rows.append(np.full(G, t)); cols.append(t * G + np.arange(G)); vals.append(-np.ones(G))
        r0 = T
        g = np.arange(G)
        for t in range(1, T):                                # ramp up / down
            up = r0 + 2 * ((t - 1) * G + g)
            rows += [up, up, up + 1, up + 1]
            cols += [t * G + g, (t - 1) * G + g, (t - 1) * G + g, t * G + g]
            vals += [np.ones(G), -np.ones(G), np.ones(G), -np.ones(G)]
        m = T + 2 * G * (T - 1)
        self.A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(m, n))
        self.A.sort_indices()
        self.m, self.n = m, n
        self.no_eq = sp.csr_matrix((0, n))
        # vectors equal for every instance are one object each (the batched engine stores them once)
        self.lb, self.ub = np.zeros(n), np.tile(self.pmax, T)
        self.b_eq = np.zeros(0)

    def params(self, k: int):
        """(demand, cost) of instance k."""
        rng = _rng(3, k)
        d = self.demand * (1.0 + rng.uniform(-0.2, 0.2, self.T))
        c = self.cost * (1.0 + rng.uniform(-0.2, 0.2, self.G))
        return d, c

    def features(self, d, c) -> np.ndarray:
        """Network input: demand and cost relative to the base instance."""
        return np.concatenate([d / self.demand - 1.0, c / self.cost - 1.0])

    def problem(self, k: int) -> Problem:
        d, c = self.params(k)
        return self.problem_from(d, c, name=f"dispatch_{k}")

    def features_of(self, ids) -> np.ndarray:
        """Network inputs of instances ``ids`` [B, T + G]."""
        return np.stack([self.features(*self.params(int(k))) for k in ids])

    def rhs(self, d) -> np.ndarray:
        G, T = self.G, self.T
        ramp = np.repeat(np.tile(self.ramp, T - 1), 2)       # [up, down] per (t, g)
        return np.concatenate([-d, ramp])

    def objective(self, c) -> np.ndarray:
        return np.tile(c, self.T)

    def problem_from(self, d, c, name: str = "dispatch") -> Problem:
        n = self.n
        # every instance references the same matrix objects (the batched engine stores them once)
        return Problem(c=self.objective(c), A_ub=self.A, b_ub=self.rhs(d), A_eq=self.no_eq,
                       b_eq=self.b_eq, lb=self.lb, ub=self.ub, integer_mask=np.zeros(n, dtype=bool), name=name)

    def params_from(self, problems):
        """(demand [B, T], cost [B, G]) read back from instances' own vectors (demand
        from the first T rows, -b_ub[:T]; costs from the first period's columns, c[:G])."""
        T, G = self.T, self.G
        d = -np.stack([np.asarray(p.b_ub[:T], float) for p in problems])
        c = np.stack([np.asarray(p.c[:G], float) for p in problems])
        return d, c

    def features_from(self, problems) -> np.ndarray:
        """Network inputs [B, T + G] read back from instances' own vectors."""
        d, c = self.params_from(problems)
        return np.concatenate([d / self.demand - 1.0, c / self.cost - 1.0], axis=1)

    # ------------------------------------------------------------ node model
    # One node per variable p_gt (node j = t*G + g, the column index).  Its input row
    # gathers, from a per-instance vector and a shared one:
    #   demand 4 d_{t+k} / sum(Pmax), k = -3..3 (clamped to the horizon);
    #   cost c_g / 60, its perturbation c_g / cost_g - 1, Pmax_g / 400, and its
    #   merit-order position: 2 (capacity strictly cheaper than g) / sum(Pmax) and
    #   2 (that + Pmax_g) / sum(Pmax);
    #   the merit-order dispatch that ignores ramps, clip(d_t' - cheaper_g, 0, Pmax_g) / Pmax_g,
    #   at t' = t-2..t+2, and its marginal price / 60 at t-1..t+1;
    #   the period t / (T-1), first / last period flags;
    #   the whole demand profile d / demand - 1.
    # Targets (4 channels): p_gt / Pmax_g, the ramp-up and ramp-down duals of (t, g) / 10
    # (rows of t >= 1), and the period's demand dual / 60.
    NODE_WINDOW_D, NODE_WINDOW_X, NODE_WINDOW_P = 3, 2, 1

    def _node_offsets(self):
        T, G = self.T, self.G
        o = dict(D=0, C=T, U=T + G, CC=T + 2 * G, CC2=T + 3 * G, P=T + 4 * G, M=2 * T + 4 * G, X=3 * T + 4 * G)
        return o, 3 * T + 4 * G + T * G

    def node_spec(self):
        """Index table, shared vector and target map of the node model (see above)."""
        from .learned.nodes import NodeSpec
        T, G = self.T, self.G
        o, L = self._node_offsets()
        shared = np.concatenate([self.pmax / 400.0, np.arange(T) / max(T - 1, 1), (np.arange(T) == 0),
                                 (np.arange(T) == T - 1)]).astype(np.float32)
        sPM, sT, sF, sL = L, L + G, L + G + T, L + G + 2 * T
        t = np.repeat(np.arange(T), G)
        g = np.tile(np.arange(G), T)
        cl = lambda a: np.clip(a, 0, T - 1)
        cols = [o["D"] + cl(t + k) for k in range(-self.NODE_WINDOW_D, self.NODE_WINDOW_D + 1)]
        cols += [o["C"] + g, o["U"] + g, o["CC"] + g, o["CC2"] + g, sPM + g]
        cols += [o["X"] + cl(t + k) * G + g for k in range(-self.NODE_WINDOW_X, self.NODE_WINDOW_X + 1)]
        cols += [o["P"] + cl(t + k) for k in range(-self.NODE_WINDOW_P, self.NODE_WINDOW_P + 1)]
        cols += [sT + t, sF + t, sL + t]
        cols += [np.full(T * G, o["M"] + k) for k in range(T)]
        table = np.stack(cols, axis=1).astype(np.int32)
        n = self.n
        up = np.where(t >= 1, n + T + 2 * ((t - 1) * G + g), -1)
        tidx = np.stack([np.arange(n), up, np.where(t >= 1, up + 1, -1), n + t], axis=1).astype(np.int32)
        tscale = np.stack([1.0 / np.tile(self.pmax, T), np.full(n, 0.1), np.full(n, 0.1), np.full(n, 1.0 / 60)],
                          axis=1).astype(np.float32)
        return NodeSpec(nodes=n, table=table, shared=shared, src_len=L, target_index=tidx, target_scale=tscale)

    def node_neighbours(self, reach: int = 2) -> np.ndarray:
        """[n, 2 reach + 1] nodes of the same generator at periods t-reach..t+reach
        (clamped to the horizon): the inputs a second node-model stage reads."""
        T, G = self.T, self.G
        t = np.repeat(np.arange(T), G)
        g = np.tile(np.arange(G), T)
        return np.stack([np.clip(t + k, 0, T - 1) * G + g for k in range(-reach, reach + 1)], axis=1).astype(np.int32)

    def node_inputs_params(self, d: np.ndarray, c: np.ndarray) -> np.ndarray:
        """Per-instance vectors [B, L] (float32) of demand [B, T] and cost [B, G]."""
        T, G = self.T, self.G
        d, c = np.atleast_2d(d), np.atleast_2d(c)
        B = d.shape[0]
        o, L = self._node_offsets()
        S = self.pmax.sum()
        order = np.argsort(c, axis=1, kind="stable")
        pm = self.pmax[order]
        cum = np.cumsum(pm, axis=1)
        cheaper = np.empty_like(c)
        np.put_along_axis(cheaper, order, cum - pm, axis=1)
        v = np.empty((B, L), np.float32)
        # [B, T, G] terms in blocks of instances (the temporaries stay in cache; same values)
        marg = np.empty((B, T), np.int64)
        for s in range(0, B, 32):
            e = min(B, s + 32)
            marg[s:e] = np.minimum((cum[s:e, None, :] < d[s:e, :, None]).sum(axis=2), G - 1)
            t = d[s:e, :, None] - cheaper[s:e, None, :]
            np.clip(t, 0.0, self.pmax, out=t)
            t /= self.pmax
            v[s:e, o["X"]:] = t.reshape(e - s, T * G)
        price = np.take_along_axis(np.take_along_axis(c, order, axis=1), marg, axis=1)
        v[:, o["D"]:o["D"] + T] = 4.0 * d / S
        v[:, o["C"]:o["C"] + G] = c / 60.0
        v[:, o["U"]:o["U"] + G] = c / self.cost - 1.0
        v[:, o["CC"]:o["CC"] + G] = 2.0 * cheaper / S
        v[:, o["CC2"]:o["CC2"] + G] = 2.0 * (cheaper + self.pmax) / S
        v[:, o["P"]:o["P"] + T] = price / 60.0
        v[:, o["M"]:o["M"] + T] = d / self.demand - 1.0
        return v

    def node_inputs_device(self, c_dev, b_dev, stream=None):
        """:meth:`node_inputs_params` on the device, bitwise, from the LPs' own cost and
        right-hand-side rows already in HBM: ``c_dev`` [B, >= G] and ``b_dev`` [B, >= T]
        FP64 device tensors (row k = instance k's ``c`` and ``b_ub``, e.g. the family
        engine's staging, :meth:`FamilyEngine.staged`).  Returns V [B, L] FP32 on
        the device, written on ``stream`` (default: the current stream)."""
        import ctypes as C
        import torch
        from .engine import _lib as L
        T, G = self.T, self.G
        B = c_dev.shape[0]
        for t_, w in ((c_dev, G), (b_dev, T)):
            if (t_.dtype != torch.float64 or t_.dim() != 2 or t_.shape[0] != B or t_.shape[1] < w
                    or t_.stride(1) != 1):
                raise ValueError("c_dev / b_dev must be [B, >= G] / [B, >= T] float64 device rows")
        dv = c_dev.device
        key = str(dv)
        if getattr(self, "_dev_consts", None) is None:
            self._dev_consts = {}
        if key not in self._dev_consts:
            self._dev_consts[key] = tuple(torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dv)
                                          for a in (self.pmax, self.cost, self.demand))
        pm, co, de = self._dev_consts[key]
        o, L_ = self._node_offsets()
        st = stream if stream is not None else torch.cuda.current_stream(dv)
        with torch.cuda.stream(st):
            V = torch.empty((B, L_), dtype=torch.float32, device=dv)
        a = L.PdlpDispatchInputs(count=B, G=G, T=T, ld_c=c_dev.stride(0), ld_b=b_dev.stride(0), ld_out=L_,
                                 off_d=o["D"], off_c=o["C"], off_u=o["U"], off_cc=o["CC"], off_cc2=o["CC2"],
                                 off_p=o["P"], off_m=o["M"], off_x=o["X"], pmax_sum=float(self.pmax.sum()),
                                 c=c_dev.data_ptr(), b=b_dev.data_ptr(), pmax=pm.data_ptr(), cost=co.data_ptr(),
                                 demand=de.data_ptr(), out=V.data_ptr())
        L.check(L.lib().pdlp_dispatch_inputs(C.byref(a), C.c_void_p(st.cuda_stream)), "pdlp_dispatch_inputs")
        return V

    def node_inputs_staged(self, problems, eng):
        """Network inputs of a batch the family engine ``eng`` has just loaded: on the
        device from its staged costs and right-hand sides (:meth:`node_inputs_device`,
        on the engine's stream), or on the host when the engine stores either once."""
        c_dev, b_dev = eng.staged("c"), eng.staged("b")
        if c_dev is None or b_dev is None:
            return self.node_inputs_from(problems)
        return self.node_inputs_device(c_dev, b_dev, eng.stream)

    def node_inputs(self, ids) -> np.ndarray:
        P = [self.params(int(k)) for k in ids]
        return self.node_inputs_params(np.stack([p[0] for p in P]), np.stack([p[1] for p in P]))

    def node_inputs_from(self, problems) -> np.ndarray:
        return self.node_inputs_params(*self.params_from(problems))


class SetCoverNodes:
    """LP relaxations of a set-covering MILP and branch-and-bound nodes.

    Base (seed [4, 0]): binary A (rows × cols) with ``round(density·rows·cols)``
    ones at distinct positions, every row covered at least twice; c ~ U[1,100];
    rows ``−A x ≤ −1``, 0 ≤ x ≤ 1 (``benchmarks/miplib_benchmark.py:133-157``
    form).  Node k (seed [4, 1+k]) fixes f ~ U{1..50} distinct variables to
    0 or 1 (lb = ub), tightening bounds only, as node problems are built in
    ``solver/milp/branch_and_bound.py:95-113``.  Fixings that leave a row with
    no free or one-fixed variable are redrawn, so every node is feasible.
    """

    def __init__(self, rows: int = 1000, cols: int = 10000, density: float = 0.005, base_seed: int = 0):
        rng = _rng(4, base_seed)
        k = int(round(density * rows * cols))
        pos = np.sort(rng.choice(rows * cols, size=k, replace=False))
        r, c = pos // cols, pos % cols
        A = sp.csr_matrix((np.ones(k), (r, c)), shape=(rows, cols)).tolil()
        cnt = np.asarray(A.sum(axis=1)).ravel()
        for i in np.flatnonzero(cnt < 2):
            extra = rng.choice(cols, size=2, replace=False)
            A[i, extra] = 1.0
        self.A = sp.csr_matrix(A)
        self.A.sum_duplicates()
        self.A.data[:] = 1.0
        self.A.sort_indices()
        self.c = rng.uniform(1.0, 100.0, cols)
        self.rows, self.cols = rows, cols
        # shared by every node problem (the batched engine stores them once)
        self.negA = sp.csr_matrix(-self.A)
        self.no_eq = sp.csr_matrix((0, cols))

    @property
    def n(self) -> int:
        return self.cols

    @property
    def m(self) -> int:
        return self.rows

    def problem(self, k: int) -> Problem:
        """Node k (the family interface of the learned warm start)."""
        return self.node(k)

    def fixings(self, k: int) -> np.ndarray:
        """+1 for a variable fixed to 1, -1 for one fixed to 0, 0 if free."""
        p = self.node(k)
        fixed = p.lb == p.ub
        return np.where(fixed, 2.0 * p.lb - 1.0, 0.0)

    def features_of(self, ids) -> np.ndarray:
        """Network inputs of nodes ``ids``: their fixings [B, cols]."""
        return np.stack([self.fixings(int(k)) for k in ids])

    def root(self) -> Problem:
        return self.node_from(np.zeros(self.cols), np.ones(self.cols), name="setcover_root")

    def node_from(self, lb, ub, name="setcover_node") -> Problem:
        return Problem(c=self.c.copy(), A_ub=self.negA, b_ub=-np.ones(self.rows), A_eq=self.no_eq,
                       b_eq=np.zeros(0), lb=np.asarray(lb, float), ub=np.asarray(ub, float),
                       integer_mask=np.zeros(self.cols, dtype=bool), name=name)

    def node(self, k: int) -> Problem:
        rng = _rng(4, 1 + k)
        for _ in range(100):
            f = int(rng.integers(1, 51))
            idx = rng.choice(self.cols, size=f, replace=False)
            val = rng.integers(0, 2, size=f).astype(float)
            lb = np.zeros(self.cols)
            ub = np.ones(self.cols)
            lb[idx] = val
            ub[idx] = val
            # every row must keep a variable that can be 1
            if np.all(self.A @ ub >= 1.0):
                return self.node_from(lb, ub, name=f"setcover_node_{k}")
        raise RuntimeError("could not draw a feasible node")


def unit_cost_set_cover(m: int = 120, n: int = 300, density: float = 0.04, seed: int = 0) -> Problem:
    """A set-covering MILP that branch-and-bound has to search: min c'x,
    A x >= 1 (as -A x <= -1), x binary, A random 0/1 with the given density
    (every row has at least one 1), costs 1 + U[0, 0.01] (near-ties: many
    optimal-looking subsets).  The covering structure of
    benchmarks/miplib_benchmark.py:133-157 with near-unit costs."""
    rng = _rng(5, seed)
    A = (rng.random((m, n)) < density).astype(float)
    A[np.arange(m), rng.integers(0, n, m)] = 1.0
    return Problem(c=1.0 + 0.01 * rng.random(n), A_ub=sp.csr_matrix(-A), b_ub=-np.ones(m), A_eq=_empty(n),
                   b_eq=np.zeros(0), lb=np.zeros(n), ub=np.ones(n), integer_mask=np.ones(n, bool),
                   name=f"setcover_unit_{m}x{n}_{seed}")
