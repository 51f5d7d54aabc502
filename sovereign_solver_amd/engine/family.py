"""Shared-matrix batched PDHG engine ("family"): B LPs with one constraint matrix.

Parametric LP families (the same K, varying costs and right-hand sides) and
branch-and-bound nodes (the same K, varying bounds) step in lockstep on the
device (``pdlp_family_*`` in ``include/pdlp_hip.h``).  Every per-LP vector is
stored in tiles of 64 LPs with the LP index fastest, so each wavefront
updates one row or column of 64 LPs with coalesced loads; vectors equal for
all LPs are stored once.  Each LP keeps its own primal weight, restart state
and termination, exactly as in the persistent batched engine.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Dict, Optional, Sequence

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib as L
from .batch import OUT_DTYPE, _STATUS
from .device import dev_tensor, ptr, require_gpu, stream_handle
from .forms import ROW_EQ, ROW_LE
from ..config import WARMSTART_FAILOVER_ITERS
from ..problem import Problem

TILE = L.PDLP_FAM_TILE
MAX_GRAPH_BLOCKS = 1


def _canonical(A) -> sp.csr_matrix:
    A = sp.csr_matrix(A, dtype=np.float64)
    A.sum_duplicates()
    A.sort_indices()
    A.indices = A.indices.astype(np.int32, copy=False)
    A.indptr = A.indptr.astype(np.int32, copy=False)
    return A


def _same_matrix(A, B) -> bool:
    if A is B:
        return True
    A, B = sp.csr_matrix(A), sp.csr_matrix(B)
    if A.shape != B.shape or A.nnz != B.nnz:
        return False
    A, B = _canonical(A), _canonical(B)
    return (np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
            and np.array_equal(A.data, B.data))


def _h2d(a: np.ndarray, dv) -> torch.Tensor:
    """A host array on the device, copied on the current stream from pinned memory
    without waiting for that stream's earlier work."""
    return torch.from_numpy(np.ascontiguousarray(a)).pin_memory().to(dv, non_blocking=True)


def to_tiles(a: np.ndarray, tiles: int) -> np.ndarray:
    """[B, L] (LP-major) -> flat tiled layout [tiles, L, 64], padded LPs zero."""
    B, Ln = a.shape
    pad = np.zeros((tiles * TILE, Ln), a.dtype)
    pad[:B] = a
    return np.ascontiguousarray(pad.reshape(tiles, TILE, Ln).transpose(0, 2, 1)).ravel()


def from_tiles(flat: np.ndarray, tiles: int, Ln: int, B: int) -> np.ndarray:
    """Inverse of :func:`to_tiles`: [B, L]."""
    return flat.reshape(tiles, Ln, TILE).transpose(0, 2, 1).reshape(tiles * TILE, Ln)[:B]


def _same_vectors(vectors) -> bool:
    """Whether every vector holds the first one's values (the same object, or equal
    entries); an unequal pair usually shows at once."""
    first = vectors[0]
    f0 = None
    for v in vectors:
        if v is first:
            continue
        if f0 is None:
            f0 = np.asarray(first, float)
        a = np.asarray(v, float)
        if a.shape != f0.shape or not np.array_equal(a, f0):
            return False
    return True


class FamilyBatch:
    """Host description of B box-form LPs that share K = [A_ub; A_eq].

    Holds the matrix analysis (device numbering, K and K^T, row kinds, the
    diagonal step sizes) and which problem vectors every LP shares (stored
    once on the device) — c, b = [b_ub; b_eq], the bounds; the per-LP vectors
    are packed on the device by :meth:`FamilyEngine.load`
    (``pdlp_family_load``).  The device numbers rows and columns so that those
    with more than ``PDLP_FAM_LONG`` entries (one workgroup each) come last:
    ``row_perm[i]`` / ``col_perm[j]`` is the problem's index of device row i /
    column j; :meth:`to_device_cols` / :meth:`from_device_cols` convert."""

    def __init__(self, problems: Sequence[Problem], long_threshold: int = L.PDLP_FAM_LONG):
        self.problems = list(problems)
        if not self.problems:
            raise ValueError("empty family")
        self.long_threshold = long_threshold
        p0 = self.problems[0]
        self.A_ub, self.A_eq = p0.A_ub, p0.A_eq
        self.m_ub, self.m_eq = p0.n_ineq, p0.n_eq
        self._check_matrices(self.problems)
        K0 = _canonical(sp.vstack([sp.csr_matrix(p0.A_ub), sp.csr_matrix(p0.A_eq)], format="csr"))
        self.K_orig = K0
        m, n = K0.shape
        if m == 0 or n == 0:
            raise ValueError("the family engine needs m > 0 and n > 0")
        self.m, self.n = m, n
        rl, cl = np.diff(K0.indptr), np.diff(K0.tocsc().indptr)
        lt = long_threshold
        self.row_perm = np.concatenate([np.flatnonzero(rl <= lt), np.flatnonzero(rl > lt)])
        self.col_perm = np.concatenate([np.flatnonzero(cl <= lt), np.flatnonzero(cl > lt)])
        self.long_rows = int((rl > lt).sum())
        self.long_cols = int((cl > lt).sum())
        K = _canonical(K0[self.row_perm][:, self.col_perm])
        self.K, self.KT = K, _canonical(K.T.tocsr())
        kind0 = np.concatenate([np.full(self.m_ub, ROW_LE, np.int8), np.full(self.m_eq, ROW_EQ, np.int8)])
        self.kind = kind0[self.row_perm]
        aK = abs(K)
        colsum = np.asarray(aK.sum(axis=0)).ravel()
        rowsum = np.asarray(aK.sum(axis=1)).ravel()
        self.tau_hat = (0.999 / np.maximum(colsum, 1e-4)).astype(np.float32)
        self.sigma_hat = (0.999 / np.maximum(rowsum, 1e-4)).astype(np.float32)
        self.row_metric = np.maximum(rowsum, 1e-4) / 0.999
        self._analyse_sharing()

    @property
    def count(self) -> int:
        return len(self.problems)

    @property
    def tiles(self) -> int:
        return (self.count + TILE - 1) // TILE

    def _check_matrices(self, problems) -> None:
        p0 = self.problems[0]
        for k, p in enumerate(problems):
            if (p.n_ineq, p.n_eq, p.n_vars) != (self.m_ub, self.m_eq, p0.n_vars):
                raise ValueError(f"LP {k}: shape differs from LP 0")
            if not (_same_matrix(p.A_ub, self.A_ub) and _same_matrix(p.A_eq, self.A_eq)):
                raise ValueError(f"LP {k}: constraint matrix differs from LP 0 (use the persistent batch engine)")

    def _pattern(self, problems) -> dict:
        """Which of c, b, bounds every problem shares."""
        return {"c": _same_vectors([p.c for p in problems]),
                "b": _same_vectors([p.b_ub for p in problems]) and _same_vectors([p.b_eq for p in problems]),
                "bounds": _same_vectors([p.lb for p in problems]) and _same_vectors([p.ub for p in problems])}

    def _analyse_sharing(self) -> None:
        pat = self._pattern(self.problems)
        self.pattern = pat
        F = L.PDLP_FAM_SHARED
        bits = 0
        if pat["c"]:
            bits |= F["c"] | F["c64"]
        if pat["b"]:
            bits |= F["b"] | F["b64"]
        if pat["bounds"]:
            bits |= F["lo"] | F["hi"] | F["lo_shift"] | F["ub_span"]
        if pat["b"] and pat["bounds"]:
            bits |= F["b_shift"]
        self.shared = bits
        # the single copies of the shared vectors, in device order
        p0, cp, rp = self.problems[0], self.col_perm, self.row_perm
        sv = {}
        if pat["c"]:
            sv["c"] = sv["c64"] = np.asarray(p0.c, float)[cp]
        if pat["bounds"]:
            lo, hi = np.asarray(p0.lb, float)[cp], np.asarray(p0.ub, float)[cp]
            lbf = np.where(np.isfinite(lo), lo, 0.0)
            sv["lo"], sv["hi"], sv["lo_shift"] = lo, hi, lbf
            sv["ub_span"] = np.where(np.isfinite(hi), hi - lbf, np.inf)
        if pat["b"]:
            sv["b"] = sv["b64"] = np.concatenate([np.asarray(p0.b_ub, float), np.asarray(p0.b_eq, float)])[rp]
        if pat["b"] and pat["bounds"]:
            sv["b_shift"] = sv["b"] - self.K @ sv["lo_shift"]
        self.shared_vec = sv

    def with_problems(self, problems: Sequence[Problem]) -> "FamilyBatch":
        """The same family (matrix analysis, shared vectors) with other LPs: they must
        hold the same matrices and share exactly what this batch shares, with equal
        values."""
        problems = list(problems)
        if len(problems) != self.count:
            raise ValueError(f"{len(problems)} LPs for an engine of {self.count}")
        self._check_matrices(problems)
        pat = self._pattern(problems)
        if pat != self.pattern:
            raise ValueError(f"the LPs share {pat}, the engine's batch {self.pattern}: build a new engine")
        q, p0 = problems[0], self.problems[0]
        eq = lambda a, b: a is b or np.array_equal(np.asarray(a, float), np.asarray(b, float))
        if ((pat["c"] and not eq(q.c, p0.c)) or (pat["b"] and not (eq(q.b_ub, p0.b_ub) and eq(q.b_eq, p0.b_eq)))
                or (pat["bounds"] and not (eq(q.lb, p0.lb) and eq(q.ub, p0.ub)))):
            raise ValueError("a vector the engine stores once differs in the new LPs: build a new engine")
        fb = object.__new__(FamilyBatch)
        fb.__dict__.update(self.__dict__)
        fb.problems = problems
        return fb

    def to_device_cols(self, X: np.ndarray) -> np.ndarray:
        return X[:, self.col_perm]

    def to_device_rows(self, Y: np.ndarray) -> np.ndarray:
        return Y[:, self.row_perm]

    def from_device_cols(self, Xd: np.ndarray) -> np.ndarray:
        X = np.empty_like(Xd)
        X[:, self.col_perm] = Xd
        return X

    def from_device_rows(self, Yd: np.ndarray) -> np.ndarray:
        Y = np.empty_like(Yd)
        Y[:, self.row_perm] = Yd
        return Y

    def is_shared(self, name: str) -> bool:
        return bool(self.shared & L.PDLP_FAM_SHARED[name])

    def bounds(self):
        """[B, n] lower and upper bounds in the problems' own column order."""
        return (np.stack([np.asarray(p.lb, float) for p in self.problems]),
                np.stack([np.asarray(p.ub, float) for p in self.problems]))

    def column_pass_bytes(self) -> float:
        """Algorithmic bytes of one column pass over all B LPs: K^T's CSR once
        (8 B per entry, 4 B per row pointer), tau_hat, y gathered once (4 B per
        row and LP), c / lo / hi (once if shared), x read and written, xbar
        written, the x partial sum read and written (4 B each per column and LP)."""
        m, n, B = self.m, self.n, self.count
        per = lambda name, L_: L_ * 4 * (1 if self.is_shared(name) else B)
        return float(8 * self.K.nnz + 4 * (n + 1) + 4 * n + 4 * m * B + per("c", n) + per("lo", n) +
                     per("hi", n) + 5 * 4 * n * B)

    def row_pass_bytes(self) -> float:
        """Algorithmic bytes of one row pass over all B LPs: K's CSR once,
        sigma_hat and the row kinds, xbar gathered once, b (once if shared), y
        read and written, the y partial sum read and written."""
        m, n, B = self.m, self.n, self.count
        per = lambda name, L_: L_ * 4 * (1 if self.is_shared(name) else B)
        return float(8 * self.K.nnz + 4 * (m + 1) + 5 * m + 4 * n * B + per("b", m) + 4 * 4 * m * B)

    def algorithmic_bytes_per_iteration(self) -> float:
        return self.column_pass_bytes() + self.row_pass_bytes()


@dataclass
class StreamBatch:
    """One batch of LPs of a stream through a :class:`FamilyEngine`
    (:meth:`FamilyEngine.run_stream`): ``src`` the LP-major FP64 device sources
    [k, len] of the per-LP vectors, named as :meth:`FamilyEngine.staged` names
    them (exactly the engine's non-shared ones); ``start`` None (cold) or a
    contiguous FP32 device matrix [k, d] read as by
    :meth:`FamilyEngine.start_from_device` (``x_off``, ``y_off``, ``w_col``),
    complete once ``ready`` (an event, or None) has passed.  ``count``: the
    number of LPs, needed only when neither sources nor a start give it."""
    src: Dict[str, torch.Tensor]
    start: Optional[torch.Tensor] = None
    ready: Optional[torch.cuda.Event] = None
    x_off: int = 0
    y_off: int = -1
    w_col: int = -1
    count: Optional[int] = None

    def __post_init__(self):
        sizes = {int(v.shape[0]) for v in self.src.values()}
        if self.start is not None:
            sizes.add(int(self.start.shape[0]))
        if self.count is not None:
            sizes.add(int(self.count))
        if len(sizes) != 1:
            raise ValueError("a stream batch needs its sources, start and count for the same LPs")
        self.count = sizes.pop()


class FamilyEngine:
    """Device state of a :class:`FamilyBatch` and the lockstep solve loop."""

    def __init__(self, fb: FamilyBatch, scheme: str = "adaptive", check_every: int = 50,
                 items_per_wave: int = 1, use_graphs: bool = True, compaction: bool = True,
                 xcd_affinity: bool = False, failover_iters: int = WARMSTART_FAILOVER_ITERS):
        flags = {"reference": 0, "adaptive": 1, "adaptive_fixed": 3, "adaptive_start": 5}
        if scheme not in flags:
            raise ValueError("scheme must be 'adaptive', 'adaptive_fixed', 'adaptive_start' or 'reference'")
        self.lib = L.lib()
        self.device = dv = require_gpu()
        self.fb = fb
        self.scheme = scheme
        self.flags = flags[scheme]
        self.check_every = check_every
        self.items_per_wave = items_per_wave
        self.use_graphs = use_graphs
        self.compaction = compaction
        self.max_graph_blocks = MAX_GRAPH_BLOCKS
        self.stream = torch.cuda.Stream(device=dv)
        f32, f64, i32 = torch.float32, torch.float64, torch.int32
        B, Tn = fb.count, fb.tiles * TILE
        n, m = fb.n, fb.m
        with torch.cuda.stream(self.stream):
            t = {
                "row_ptr": dev_tensor(fb.K.indptr, i32, dv), "col_idx": dev_tensor(fb.K.indices, i32, dv),
                "val": dev_tensor(fb.K.data, f32, dv),
                "t_row_ptr": dev_tensor(fb.KT.indptr, i32, dv), "t_col_idx": dev_tensor(fb.KT.indices, i32, dv),
                "t_val": dev_tensor(fb.KT.data, f32, dv),
                "row_kind": dev_tensor(fb.kind, torch.int8, dv), "tau_hat": dev_tensor(fb.tau_hat, f32, dv),
                "sigma_hat": dev_tensor(fb.sigma_hat, f32, dv), "row_metric": dev_tensor(fb.row_metric, f64, dv),
                "row_orig": dev_tensor(fb.row_perm.astype(np.int32), i32, dv),
                "col_orig": dev_tensor(fb.col_perm.astype(np.int32), i32, dv),
                "row_dev": dev_tensor(np.argsort(fb.row_perm).astype(np.int32), i32, dv),
                "col_dev": dev_tensor(np.argsort(fb.col_perm).astype(np.int32), i32, dv),
            }
            for name, ln in (("x", n), ("xbar", n), ("x_psum", n), ("x_last", n),
                             ("y", m), ("y_psum", m), ("y_last", m)):
                t[name] = torch.zeros(Tn * ln, dtype=f32, device=dv)
            t["x_sum"] = torch.zeros(Tn * n, dtype=f64, device=dv)
            t["y_sum"] = torch.zeros(Tn * m, dtype=f64, device=dv)
            t["lp"] = torch.zeros(B * C.sizeof(L.PdlpFamilyLP), dtype=torch.uint8, device=dv)
            t["out"] = torch.zeros(B * OUT_DTYPE.itemsize, dtype=torch.uint8, device=dv)
            t["run_mask"] = torch.zeros(fb.tiles, dtype=torch.int64, device=dv)
            t["running"] = torch.zeros(1, dtype=i32, device=dv)
            t["running_sub"] = torch.zeros(1, dtype=i32, device=dv)     # the count of a view's own check
            t["X_out"] = torch.zeros((B, n), dtype=f64, device=dv)
            t["Y_out"] = torch.zeros((B, m), dtype=f64, device=dv)
            t["out_final"] = torch.zeros(B * OUT_DTYPE.itemsize, dtype=torch.uint8, device=dv)
            t["partials"] = torch.zeros(2 * L.PDLP_FAM_KKT_CHUNKS * L.PDLP_FAM_NSUM * Tn, dtype=f64, device=dv)
            # problem vectors: one copy of the shared ones, tiled per-LP arrays (filled by load)
            # with a copy each that restores the slot order after compaction
            self._orig = {}
            for name, dt in self._DATA.items():
                ln = n if name in self._COLS else m
                if fb.is_shared(name):
                    t[name] = dev_tensor(fb.shared_vec[name], dt, dv)
                else:
                    t[name] = torch.zeros(Tn * ln, dtype=dt, device=dv)
                    self._orig[name] = torch.zeros(Tn * ln, dtype=dt, device=dv)
            # LP-major FP64 staging of the per-LP sources (pinned host copy, device copy)
            self._src = {}
            for name, ln, per_lp in (("c", n, not fb.is_shared("c")), ("b", m, not fb.is_shared("b")),
                                     ("lo", n, not fb.is_shared("lo")), ("hi", n, not fb.is_shared("hi"))):
                if per_lp:
                    self._src[name] = (torch.empty((B, ln), dtype=f64).pin_memory(),
                                       torch.empty((B, ln), dtype=f64, device=dv))
            t["norms"] = torch.zeros(3 * B, dtype=f64, device=dv)
            self.norms_host = torch.zeros(3 * B, dtype=f64).pin_memory()
        self.t = t
        self.stream.synchronize()
        # per-LP tiled arrays that compaction moves (window partial sums and xbar
        # are zero / rewritten at a check boundary and need not move)
        self._per_lp = [("x", n), ("x_last", n), ("x_sum", n), ("y", m), ("y_last", m), ("y_sum", m)]
        self._per_lp += [(name, n if name in ("c", "lo", "hi", "c64", "lo_shift", "ub_span") else m)
                         for name in self._DATA if not fb.is_shared(name)]
        self.running_host = torch.zeros((2, 1), dtype=i32).pin_memory()
        self.mask_host = torch.zeros((2, fb.tiles), dtype=torch.int64).pin_memory()
        self.compactions = 0
        self.args = L.PdlpFamily(m=m, n=n, count=B, tiles=fb.tiles, check_every=check_every, flags=self.flags,
                                 shared=fb.shared, items_per_wave=items_per_wave, long_rows=fb.long_rows,
                                 long_cols=fb.long_cols, xcd_affinity=int(xcd_affinity),
                                 failover_iters=int(failover_iters), tol=0.0)
        first = [name for name, _ in L.PdlpFamily._fields_].index("row_ptr")
        for name, _ in L.PdlpFamily._fields_[first:]:
            setattr(self.args, name, ptr(t[name]))
        self._graphs = {}
        self._cap_stream = torch.cuda.Stream(device=dv)   # graphs are captured here, off the busy stream
        self._host = None
        self.load()

    _DATA = {"c": torch.float32, "lo": torch.float32, "hi": torch.float32, "b": torch.float32,
             "c64": torch.float64, "lo_shift": torch.float64, "ub_span": torch.float64, "b64": torch.float64,
             "b_shift": torch.float64}
    _COLS = ("c", "lo", "hi", "c64", "lo_shift", "ub_span")

    def host_sources(self, problems) -> Dict[str, torch.Tensor]:
        """Fresh pinned LP-major FP64 copies [len(problems), len] of the per-LP vectors
        this engine does not store once (``c``, ``b`` = [b_ub, b_eq], ``lo``, ``hi``),
        e.g. to upload as a :class:`StreamBatch`'s sources."""
        fb = self.fb
        out = {name: torch.empty((len(problems), fb.n if name != "b" else fb.m), dtype=torch.float64).pin_memory()
               for name in self._src}
        self._gather(problems, out)
        return out

    def _gather(self, problems, dst: Optional[Dict[str, torch.Tensor]] = None) -> None:
        """Per-LP sources into the pinned LP-major staging (or into ``dst``): one
        NumPy view of each staging array, then one row copy per LP and vector (one
        thread: the copies are short and threads only contend for the interpreter)."""
        src = {k: v[0] for k, v in self._src.items()} if dst is None else dst
        if not src:
            return
        m_ub = self.fb.m_ub
        C_ = src["c"].numpy() if "c" in src else None
        B_ = src["b"].numpy() if "b" in src else None
        LO = src["lo"].numpy() if "lo" in src else None
        HI = src["hi"].numpy() if "lo" in src else None
        for k, p in enumerate(problems):
            if C_ is not None:
                C_[k] = p.c
            if B_ is not None:
                B_[k, :m_ub] = p.b_ub
                B_[k, m_ub:] = p.b_eq
            if LO is not None:
                LO[k] = p.lb
                HI[k] = p.ub

    def load(self, problems: Optional[Sequence[Problem]] = None) -> None:
        """Problem data of the batch's LPs (or of ``problems``: the same family, count and
        sharing, see :meth:`FamilyBatch.with_problems`) onto the device: per-LP
        vectors gathered into pinned LP-major staging, copied over, and tiled with
        their derived vectors and normalisers by ``pdlp_family_load``."""
        if problems is not None:
            self.fb = self.fb.with_problems(problems)
        fb = self.fb
        self._gather(fb.problems)
        src = self._src
        with torch.cuda.stream(self.stream):
            for h, d in src.values():
                d.copy_(h, non_blocking=True)
            dp = lambda name: src[name][1].data_ptr() if name in src else None
            if self.compactions or self.args.tiles != fb.tiles:
                self.args.tiles, self.args.count = fb.tiles, fb.count
                self.compactions = 0
            L.check(self.lib.pdlp_family_load(C.byref(self.args), dp("c"), dp("b"), dp("lo"), dp("hi"),
                                              self.t["norms"].data_ptr(), self.sh), "pdlp_family_load")
            for name, orig in self._orig.items():
                orig.copy_(self.t[name])
            self.norms_host.copy_(self.t["norms"], non_blocking=True)
        self.stream.synchronize()
        nv = self.norms_host.numpy().reshape(-1, 3)
        self.norm_b, self.norm_c, self.init_weight = nv[:, 0].copy(), nv[:, 1].copy(), nv[:, 2].copy()

    def staged(self, name: str) -> Optional[torch.Tensor]:
        """The device copy [B, len] FP64 of the per-LP source ``name`` (``"c"``, ``"b"``,
        ``"lo"``, ``"hi"``) as the last :meth:`load` staged it, row k = LP k in its own
        order (``b`` = [b_ub, b_eq]); None when the batch stores that vector once."""
        return self._src[name][1] if name in self._src else None

    def _restore_data(self) -> None:
        """Per-LP problem vectors back in the original slot order (after compaction)."""
        with torch.cuda.stream(self.stream):
            for name, orig in self._orig.items():
                self.t[name].copy_(orig)

    @property
    def sh(self):
        return stream_handle(self.stream)

    def drop_graphs(self) -> None:
        for g in self._graphs.values():
            self.lib.pdlp_graph_destroy(g)
        self._graphs = {}

    def __del__(self):
        try:
            self.drop_graphs()
        except Exception:
            pass

    MAX_GRAPHS = 24

    def _graph(self, blocks: int):
        """The graph of ``blocks`` blocks for the current argument block (which it holds by
        value), captured once per distinct argument block and kept, least recently used
        dropped first, so that the tile counts a solve passes through are reused by the next."""
        key = (blocks, C.string_at(C.addressof(self.args), C.sizeof(self.args)))
        g = self._graphs.pop(key, None)
        if g is None:
            h = C.c_void_p()
            L.check(self.lib.pdlp_family_graph_capture(C.byref(self.args), blocks,
                                                       stream_handle(self._cap_stream), C.byref(h)),
                    "pdlp_family_graph_capture")
            g = h
            if len(self._graphs) >= self.MAX_GRAPHS:
                self.stream.synchronize()              # the oldest may still be running
                old = next(iter(self._graphs))
                self.lib.pdlp_graph_destroy(self._graphs.pop(old))
        self._graphs[key] = g
        return g

    def set_tol(self, tol: float) -> None:
        if self.args.tol != tol:
            self.args.tol = tol
            self.drop_graphs()          # the graphs hold the argument block by value

    # ----------------------------------------------------------------- solve
    LP_DTYPE = np.dtype([(n_, "<i4") for n_ in ("status", "iter", "window", "restarts", "checks", "done_iter",
                                                "last_restart", "action")] +
                        [(n_, "<f8") for n_ in ("norm_b", "norm_c", "ref_err", "err_prev", "w_div")] +
                        [("weight", "<f4"), ("weight0", "<f4"), ("weight_cold", "<f4"), ("warm", "<i4"),
                         ("iter0", "<i4"), ("reserved", "<i4"), ("start_err", "<f8"), ("cold_err", "<f8")])

    def _reset(self, tol: float, weight: Optional[np.ndarray], warm=False) -> None:
        """Slot bookkeeping and per-LP records for a new solve (original slot order).
        ``warm`` (one value, or one per LP): bit 0 the LP starts from a caller's point or
        weight, which the warm-start failover judges; bit 1 that start sets y and is compared
        with the cold start at init."""
        fb = self.fb
        self.set_tol(tol)
        if self.compactions or self.args.tiles != fb.tiles:
            self._restore_data()
            self.args.tiles, self.args.count = fb.tiles, fb.count
            self.compactions = 0
        self.slot_lp = np.concatenate([np.arange(fb.count), np.full(fb.tiles * TILE - fb.count, -1)])
        self.saved = np.zeros(fb.count, bool)
        w_cold = self.init_weight.copy() if self.flags & 1 else np.ones(fb.count)
        w = w_cold
        if weight is not None:
            ww = np.asarray(weight, float)
            w = np.where(np.isfinite(ww) & (ww > 0), ww, w_cold)
        lp = np.zeros(fb.count, dtype=self.LP_DTYPE)
        assert lp.dtype.itemsize == C.sizeof(L.PdlpFamilyLP)
        lp["norm_b"], lp["norm_c"], lp["weight"], lp["weight0"] = self.norm_b, self.norm_c, w, w
        lp["weight_cold"] = w_cold
        lp["warm"] = np.broadcast_to(np.asarray(warm, np.int32), (fb.count,))
        with torch.cuda.stream(self.stream):
            self.t["lp"].copy_(torch.from_numpy(lp.view(np.uint8).copy()).to(self.device))

    def start(self, tol: float, x0: Optional[np.ndarray] = None, y0: Optional[np.ndarray] = None,
              weight: Optional[np.ndarray] = None) -> None:
        """Starting point ([B, n] / [B, m], None = cold start: x = clip(0, lo, hi),
        y = 0) and per-LP primal weights (None or NaN entries: 1 for the reference
        scheme, the norm ratio for the adaptive ones)."""
        fb = self.fb
        # an LP starts warm when the caller gives it a point other than the cold start or a
        # weight (bit 0), and is judged against the cold start at init when it sets y (bit 1)
        warm = np.zeros(fb.count, np.int32)
        if weight is not None:
            warm |= np.isfinite(np.asarray(weight, float)).astype(np.int32)
        LB = UB = None
        if x0 is not None:
            LB, UB = fb.bounds()
            x0 = np.clip(np.asarray(x0, float).reshape(fb.count, fb.n), LB, UB)
            warm |= (x0 != np.clip(0.0, LB, UB)).any(axis=1).astype(np.int32)
        if y0 is not None:
            warm |= 3 * (np.asarray(y0, float).reshape(fb.count, fb.m) != 0.0).any(axis=1).astype(np.int32)
        self._reset(tol, weight, warm=warm)
        dv = self.device
        if x0 is None and y0 is None:           # cold start written on the device
            L.check(self.lib.pdlp_family_set_start(C.byref(self.args), None, 0, -1, -1, -1, None, None, self.sh),
                    "pdlp_family_set_start")
        else:
            if x0 is None:
                LB, UB = fb.bounds()
                x0 = np.clip(0.0, LB, UB)
            x = fb.to_device_cols(x0)
            if y0 is None:
                y = np.zeros((fb.count, fb.m))
            else:
                y = np.array(y0, float).reshape(fb.count, fb.m)
                y[:, : fb.m_ub] = np.minimum(y[:, : fb.m_ub], 0.0)      # y <= 0 on the <= rows
                y = fb.to_device_rows(y)
            with torch.cuda.stream(self.stream):
                self.t["x"].copy_(torch.from_numpy(to_tiles(x.astype(np.float32), fb.tiles)).to(dv))
                self.t["y"].copy_(torch.from_numpy(to_tiles(y.astype(np.float32), fb.tiles)).to(dv))
        L.check(self.lib.pdlp_family_init(C.byref(self.args), self.sh), "pdlp_family_init")

    def start_from_device(self, tol: float, src: torch.Tensor, x_off: int = 0, y_off: int = -1, w_col: int = -1,
                          scale: Optional[torch.Tensor] = None, shift: Optional[torch.Tensor] = None,
                          ready: Optional[torch.cuda.Event] = None) -> None:
        """Starting points read on the device from a row-major [B, d] FP32 matrix
        (``pdlp_family_set_start``): x from columns x_off.., y from y_off.. (-1:
        cold), the primal weight exp(.) of column w_col (-1: default), each
        column de-normalised as v * scale + shift.  ``ready``: an event the
        engine stream waits for (the producer's stream)."""
        fb = self.fb
        if src.dtype != torch.float32 or src.dim() != 2 or src.shape[0] != fb.count or not src.is_contiguous():
            raise ValueError("src must be a contiguous [count, d] float32 device tensor")
        d = src.shape[1]
        for t_ in (scale, shift):
            if t_ is not None and (t_.dtype != torch.float32 or t_.numel() != d or not t_.is_contiguous()):
                raise ValueError("scale / shift must be contiguous float32 vectors of length d")
        self._reset(tol, None, warm=(1 if x_off >= 0 or y_off >= 0 or w_col >= 0 else 0) | (2 if y_off >= 0 else 0))
        if ready is not None:
            self.stream.wait_event(ready)
        else:
            self.stream.wait_stream(torch.cuda.current_stream())
        L.check(self.lib.pdlp_family_set_start(C.byref(self.args), src.data_ptr(), d, x_off, y_off, w_col,
                                               ptr(scale), ptr(shift), self.sh), "pdlp_family_set_start")
        for t_ in (src, scale, shift):          # read on this stream: not reused before it has been
            if t_ is not None:
                t_.record_stream(self.stream)
        L.check(self.lib.pdlp_family_init(C.byref(self.args), self.sh), "pdlp_family_init")

    def run(self, max_iters: int) -> int:
        """Enqueue blocks of check_every iterations until every LP has finished or
        ``max_iters`` is reached; returns the iterations executed.

        The running-LP count and tile masks come back one block behind, so the
        GPU always has the next block queued.  When finished LPs leave enough
        lanes idle (the running LPs would fit in at most 3/4 of the live
        tiles), the engine waits, saves the finished LPs' results and packs the
        running LPs into the leading tiles (:meth:`compact`)."""
        ce = self.check_every
        done, blocks, slot = 0, 1, 0
        inflight = []
        finished = False
        while True:
            left = (max_iters - done) // ce
            if left == 0:
                break
            g = min(blocks, left)
            with torch.cuda.stream(self.stream):
                if self.use_graphs:
                    g = 1 << (g.bit_length() - 1)
                    L.check(self.lib.pdlp_graph_launch(self._graph(g), self.sh), "pdlp_graph_launch")
                else:
                    L.check(self.lib.pdlp_family_run(C.byref(self.args), g, self.sh), "pdlp_family_run")
                self.running_host[slot].copy_(self.t["running"], non_blocking=True)
                self.mask_host[slot, : self.args.tiles].copy_(self.t["run_mask"][: self.args.tiles],
                                                              non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(self.stream)
            done += g * ce
            inflight.append((slot, ev, self.args.tiles))
            slot ^= 1
            blocks = min(blocks * 2, self.max_graph_blocks)
            if len(inflight) == 2:
                s0, e0, t0 = inflight.pop(0)
                e0.synchronize()
                running = int(self.running_host[s0, 0])
                if running == 0:
                    finished = True
                    break
                live = int(np.count_nonzero(self.mask_host[s0, :t0].numpy()))
                need = (running + TILE - 1) // TILE
                if self.compaction and live >= 2 and need <= 0.75 * live:
                    self.stream.synchronize()
                    inflight.clear()
                    self.compact()
        tail = 0
        if not finished:
            tail = max_iters - done
            if tail > 0:
                L.check(self.lib.pdlp_family_iterate(C.byref(self.args), tail, self.sh), "pdlp_family_iterate")
        L.check(self.lib.pdlp_family_finish(C.byref(self.args), tail, self.sh), "pdlp_family_finish")
        return done + tail

    # ---------------------------------------------------------------- streams
    def _tiled(self):
        """(name, length, element bytes) of every per-LP tiled array."""
        fb = self.fb
        out = [("x", fb.n, 4), ("xbar", fb.n, 4), ("x_psum", fb.n, 4), ("x_last", fb.n, 4),
               ("y", fb.m, 4), ("y_psum", fb.m, 4), ("y_last", fb.m, 4), ("x_sum", fb.n, 8), ("y_sum", fb.m, 8)]
        for name, dt in self._DATA.items():
            if not fb.is_shared(name):
                out.append((name, fb.n if name in self._COLS else fb.m, 8 if dt == torch.float64 else 4))
        return out

    def _view(self, t0: int, count: int) -> L.PdlpFamily:
        """Argument block of ``count`` LPs in the slots from tile ``t0`` on: every per-LP
        pointer advanced to that tile, the view's running count in its own word."""
        a = L.PdlpFamily.from_buffer_copy(self.args)
        a.count, a.tiles = count, (count + TILE - 1) // TILE
        for name, ln, es in self._tiled():
            setattr(a, name, self.t[name].data_ptr() + t0 * TILE * ln * es)
        a.lp = self.t["lp"].data_ptr() + t0 * TILE * C.sizeof(L.PdlpFamilyLP)
        a.out = self.t["out"].data_ptr() + t0 * TILE * OUT_DTYPE.itemsize
        a.run_mask = self.t["run_mask"].data_ptr() + t0 * 8
        a.running = self.t["running_sub"].data_ptr()
        return a

    def _admit(self, t0: int, b: StreamBatch, first: int, k: int, lp0: int) -> None:
        """Rows first .. first + k of stream batch ``b`` (stream LPs lp0 ..) into the
        slots from tile t0: their data, records, start and initial check, on the
        device (``pdlp_family_load`` / ``_records`` / ``_set_start`` / ``_init`` on a
        view of those tiles), while the LPs of the other tiles keep their state."""
        v = self._view(t0, k)
        for t_ in list(b.src.values()) + ([b.start] if b.start is not None else []):
            t_.record_stream(self.stream)               # read on this stream: not reused before it has been
        norms = self.t["norms"].data_ptr() + t0 * TILE * 3 * 8
        row = lambda name: (b.src[name][first:].data_ptr() if name in b.src else None)
        L.check(self.lib.pdlp_family_load(C.byref(v), row("c"), row("b"), row("lo"), row("hi"), norms, self.sh),
                "pdlp_family_load")
        warm = 0
        if b.start is not None:
            warm = (1 if b.x_off >= 0 or b.y_off >= 0 or b.w_col >= 0 else 0) | (2 if b.y_off >= 0 else 0)
        L.check(self.lib.pdlp_family_records(C.byref(v), norms, warm, self.sh), "pdlp_family_records")
        if b.start is None:
            L.check(self.lib.pdlp_family_set_start(C.byref(v), None, 0, -1, -1, -1, None, None, self.sh),
                    "pdlp_family_set_start")
        else:
            if b.ready is not None:
                self.stream.wait_event(b.ready)
            d = b.start.shape[1]
            L.check(self.lib.pdlp_family_set_start(C.byref(v), b.start[first:].data_ptr(), d, b.x_off, b.y_off,
                                                   b.w_col, None, None, self.sh), "pdlp_family_set_start")
        L.check(self.lib.pdlp_family_init(C.byref(v), self.sh), "pdlp_family_init")
        end = t0 * TILE + k
        pad = -(-end // TILE) * TILE
        if self.slot_lp.size < pad:
            self.slot_lp = np.concatenate([self.slot_lp, np.full(pad - self.slot_lp.size, -1)])
        self.slot_lp[t0 * TILE: pad] = -1
        self.slot_lp[t0 * TILE: end] = np.arange(lp0, lp0 + k)

    def run_stream(self, tol: float, max_iters: int, batches, admit_tiles: int = 0, total: Optional[int] = None):
        """Solve a stream of batches of LPs of this engine's family through its slots:
        LPs enter free tiles as earlier ones finish (at a check: the running LPs
        are packed into the leading tiles, the next LPs of the stream loaded into
        the tiles behind them), so the device runs full tiles until the stream
        runs dry instead of waiting, batch by batch, on each batch's slowest LPs.

        Each LP runs exactly the iterations, checks and decisions it would run in
        a batch solve from the same start (its checks fall every check_every of
        its own iterations; ``max_iters``, a multiple of check_every, is its own
        limit), so its result is the same bit for bit.  ``admit_tiles``: free whole
        tiles that trigger an admission (0: an eighth of the engine's tiles).
        ``batches``: a sequence of :class:`StreamBatch`, or any iterable of them
        (read only as LPs are admitted, so its batches may be prepared while
        earlier ones run) together with ``total``, the number of LPs it yields.
        Returns (out records [N] (uint8 view of OUT_DTYPE), X [N, n], Y [N, m]) on
        the device, LPs in stream order (problems' own row and column order);
        the engine's next solve reloads its own batch."""
        fb = self.fb
        ce = self.check_every
        if max_iters <= 0 or max_iters % ce:
            raise ValueError("a stream's max_iters must be a positive multiple of check_every")
        need_src = set(self._src)

        def check(b: StreamBatch) -> StreamBatch:
            if set(b.src) != need_src:
                raise ValueError(f"stream batch sources must be exactly {sorted(need_src)}")
            for name, v in b.src.items():
                ln = fb.n if name in ("c", "lo", "hi") else fb.m
                if v.dtype != torch.float64 or v.dim() != 2 or v.shape[1] != ln or not v.is_contiguous():
                    raise ValueError(f"source {name} must be a contiguous [k, {ln}] float64 device tensor")
            if b.start is not None and (b.start.dtype != torch.float32 or b.start.dim() != 2
                                        or not b.start.is_contiguous()):
                raise ValueError("a stream batch's start must be a contiguous [k, d] float32 device tensor")
            return b

        if total is None:
            batches = [check(b) for b in batches]
            total = sum(b.count for b in batches)
        source = iter(batches)
        B, tiles = fb.count, fb.tiles
        admit_tiles = admit_tiles or max(1, tiles // 8)
        dv = self.device
        keep = (self.t["X_out"], self.t["Y_out"], self.t["out_final"])
        with torch.cuda.stream(self.stream):
            X = torch.zeros((total, fb.n), dtype=torch.float64, device=dv)
            Y = torch.zeros((total, fb.m), dtype=torch.float64, device=dv)
            O = torch.zeros(total * OUT_DTYPE.itemsize, dtype=torch.uint8, device=dv)
        self.t["X_out"], self.t["Y_out"], self.t["out_final"] = X, Y, O
        self.set_tol(tol)
        self.args.max_iters = int(max_iters)
        self.compactions += 1             # the slots' data is the stream's from here: reload on the next solve
        self.saved = np.zeros(total, bool)
        self.slot_lp = np.full(tiles * TILE, -1)
        queue = []
        more = [True]                                  # the source may still yield batches
        lp_next = 0

        def pull() -> None:
            """The next batch with LPs onto the queue (waits for a lazy source)."""
            while more[0] and not queue:
                b = next(source, None)
                if b is None:
                    more[0] = False
                elif check(b).count:
                    queue.append((b, 0))

        def fill(t0: int) -> int:
            """Admit the next LPs into tiles t0 ..; returns the slots now in use."""
            nonlocal lp_next
            slot = t0 * TILE
            while slot < B:
                pull()
                if not queue:
                    break
                b, first = queue[0]
                k = min(b.count - first, B - slot)
                if lp_next + k > total:
                    raise ValueError(f"the stream yields more than total = {total} LPs")
                self._admit(slot // TILE, b, first, k, lp_next)
                lp_next += k
                if first + k == b.count:
                    queue.pop(0)
                else:
                    queue[0] = (b, first + k)
                slot += -(-k // TILE) * TILE          # a batch's last LPs may leave a tile part empty
            return min(slot, B) if slot > t0 * TILE else t0 * TILE

        try:
            used = fill(0)
            self.args.count, self.args.tiles = used, -(-used // TILE)
            inflight, slot = [], 0
            while self.args.count > 0:
                with torch.cuda.stream(self.stream):
                    L.check(self.lib.pdlp_graph_launch(self._graph(1), self.sh), "pdlp_graph_launch")
                    self.running_host[slot].copy_(self.t["running"], non_blocking=True)
                    self.mask_host[slot, : self.args.tiles].copy_(self.t["run_mask"][: self.args.tiles],
                                                                  non_blocking=True)
                    ev = torch.cuda.Event()
                    ev.record(self.stream)
                inflight.append((slot, ev, self.args.tiles))
                slot ^= 1
                if len(inflight) < 2:
                    continue
                s0, e0, t0 = inflight.pop(0)
                e0.synchronize()
                running = int(self.running_host[s0, 0])
                need = -(-running // TILE)
                if tiles - need >= admit_tiles or running == 0:
                    pull()                             # a lazy source's next batch, while blocks run
                if queue and (tiles - need >= admit_tiles or running == 0):
                    self.stream.synchronize()
                    inflight.clear()
                    self.compact()                     # running LPs to the front, finished ones saved
                    t_run = self.args.tiles
                    used = fill(t_run)
                    if used > t_run * TILE:
                        self.args.count, self.args.tiles = used, -(-used // TILE)
                elif not queue and not more[0]:
                    if running == 0:
                        break
                    live = int(np.count_nonzero(self.mask_host[s0, :t0].numpy()))
                    if self.compaction and live >= 2 and need <= 0.75 * live:
                        self.stream.synchronize()
                        inflight.clear()
                        self.compact()
            self.collect()
            self.stream.synchronize()
        finally:
            self.t["X_out"], self.t["Y_out"], self.t["out_final"] = keep
            self.args.max_iters = 0
        return O, X, Y

    # ------------------------------------------------------------ compaction
    def _slot_index(self, perm: torch.Tensor, ln: int, new_tiles: int) -> torch.Tensor:
        """Flat source index of every element of the first ``new_tiles`` tiles after
        the slot permutation ``perm`` (new slot -> old slot)."""
        base = (perm // TILE) * (ln * TILE) + perm % TILE                     # [new_tiles*64]
        d = torch.arange(ln, device=perm.device, dtype=torch.int64) * TILE
        return (base.view(new_tiles, 1, TILE) + d.view(1, ln, 1)).reshape(-1)

    def _save(self, slots: np.ndarray) -> None:
        """Copy the results of the LPs in ``slots`` (current layout) into the
        LP-major device outputs (read back once, by :meth:`results`)."""
        if slots.size == 0:
            return
        fb = self.fb
        lps = self.slot_lp[slots]
        dv = self.device
        # every temporary is allocated on the engine stream that uses it, so the
        # caching allocator cannot hand its memory to other work before that use
        with torch.cuda.stream(self.stream):
            sl = _h2d(slots.astype(np.int64), dv)
            li = _h2d(lps.astype(np.int64), dv)
            for name, ln, dst, inv in (("x_sum", fb.n, "X_out", "col_dev"), ("y_sum", fb.m, "Y_out", "row_dev")):
                # element (LP, problem index j) sits at device index inv[j]
                base = (sl // TILE) * (ln * TILE) + sl % TILE
                d = self.t[inv].to(torch.int64) * TILE
                flat = (base.view(-1, 1) + d.view(1, ln)).reshape(-1)
                self.t[dst].index_copy_(0, li, torch.take(self.t[name], flat).view(-1, ln))
            rec = self.t["out"].view(-1, OUT_DTYPE.itemsize).index_select(0, sl)
            self.t["out_final"].view(-1, OUT_DTYPE.itemsize).index_copy_(0, li, rec)
        self.saved[lps] = True

    def compact(self) -> None:
        """Save the finished LPs of the live tiles and move the running LPs to the
        first ceil(running / 64) tiles (stream drained, at a check boundary)."""
        fb = self.fb
        T, cnt = self.args.tiles, self.args.count
        with torch.cuda.stream(self.stream):
            mask = self.t["run_mask"][:T].cpu().numpy().view(np.uint64)
        bits = ((mask[:, None] >> np.arange(TILE, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(bool).ravel()
        slots = np.arange(T * TILE)
        lp_of = np.full(T * TILE, -1)
        have = min(T * TILE, self.slot_lp.size)
        lp_of[:have] = self.slot_lp[:have]
        valid = (slots < cnt) & (lp_of >= 0)           # a stream may leave empty slots inside count
        run = bits & valid
        done_slots = slots[valid & ~run & ~self.saved[np.maximum(lp_of, 0)]]
        self._save(done_slots)
        keep = slots[run]
        R = int(keep.size)
        new_T = (R + TILE - 1) // TILE
        perm_np = np.concatenate([keep, np.full(new_T * TILE - R, keep[0] if R else 0)]).astype(np.int64)
        dv = self.device
        with torch.cuda.stream(self.stream):
            perm = _h2d(perm_np, dv)
            idx = {fb.n: self._slot_index(perm, fb.n, new_T), fb.m: self._slot_index(perm, fb.m, new_T)}
            for name, ln in self._per_lp:
                t = self.t[name]
                moved = torch.take(t, idx[ln])
                t[: moved.numel()].copy_(moved)
            for name, size in (("lp", C.sizeof(L.PdlpFamilyLP)), ("out", OUT_DTYPE.itemsize)):
                v = self.t[name].view(-1, size)
                moved = v.index_select(0, perm[:R])
                v[:R].copy_(moved)
            new_mask = np.zeros(new_T, np.uint64)
            for t_ in range(new_T):
                lanes = min(TILE, R - t_ * TILE)
                new_mask[t_] = np.uint64((1 << lanes) - 1) if lanes < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
            self.t["run_mask"][:new_T].copy_(_h2d(new_mask.view(np.int64), dv))
        self.slot_lp = np.concatenate([lp_of[perm_np[:R]], np.full(new_T * TILE - R, -1)])
        self.args.tiles = new_T
        self.args.count = R
        self.compactions += 1

    def collect(self) -> None:
        """Gather every LP's result into the LP-major device outputs (``X_out``, ``Y_out``,
        ``out_final``; problems' own order), leaving them in HBM."""
        self.stream.synchronize()
        lp_of = self.slot_lp[: self.args.count]
        has = np.flatnonzero(lp_of >= 0)           # a stream may leave empty slots
        self._save(has[~self.saved[lp_of[has]]])

    def results(self):
        """(out records, X [B, n] FP64 solutions, Y [B, m]) by LP, in the problems'
        own row and column order.  X and Y are views of pinned host buffers that
        the next solve on this engine overwrites: copy what must outlive it."""
        self.collect()
        if self._host is None:                  # pinned staging, allocated once per engine
            self._host = {k: torch.empty(self.t[k].shape, dtype=self.t[k].dtype).pin_memory()
                          for k in ("X_out", "Y_out", "out_final")}
        with torch.cuda.stream(self.stream):
            for k, h in self._host.items():
                h.copy_(self.t[k], non_blocking=True)
        self.stream.synchronize()
        out = np.frombuffer(self._host["out_final"].numpy().tobytes(), dtype=OUT_DTYPE).copy()
        return out, self._host["X_out"].numpy(), self._host["Y_out"].numpy()

    def objectives(self, X: np.ndarray) -> np.ndarray:
        """Objective of each LP at X [B, n] (problems' own column order) in each problem's
        own sense, from the engine's copy of c (the staging of per-LP costs, or the
        shared vector)."""
        if "c" in self._src:
            obj = np.einsum("ij,ij->i", self._src["c"][0].numpy(), X)
        else:
            obj = X @ np.asarray(self.fb.problems[0].c, float)
        sign = np.array([-1.0 if p.sense == "max" else 1.0 for p in self.fb.problems])
        return obj * sign

    def solve(self, tol: float, max_iters: int, x0=None, y0=None, weight=None):
        self.start(tol, x0, y0, weight)
        self.run(max_iters)
        return self.results()

    def unpack(self, out, X, Y):
        return [(_STATUS[int(out["status"][k])], X[k], Y[k], out[k]) for k in range(self.fb.count)]

    # ---------------------------------------------------------------- tuning
    def set_check_every(self, ce: int) -> None:
        """Iterations between restart checks (the captured graphs are rebuilt)."""
        if int(ce) <= 0:
            raise ValueError("check_every must be positive")
        self.check_every = int(ce)
        self.args.check_every = int(ce)
        self.drop_graphs()

    def autotune(self, tol: float = 1e-4, reps: int = 10) -> dict:
        """Pick rows / columns per wavefront (1, 2) and XCD affinity (off, on) by
        timing the column and row passes (HIP graphs, events) at the cold-start
        state with every LP running; keeps the fastest sum.  Placement changes
        no arithmetic, so results do not depend on the choice."""
        best = None
        for ipw in (1, 2):
            for aff in (0, 1):
                self.args.items_per_wave, self.args.xcd_affinity = ipw, aff
                self.start(tol)
                t = self.time_step("primal", reps) + self.time_step("dual", reps)
                if best is None or t < best[0]:
                    best = (t, ipw, aff)
        self.args.items_per_wave, self.args.xcd_affinity = best[1], best[2]
        self.items_per_wave = best[1]
        self.drop_graphs()
        self.stream.synchronize()
        return {"items_per_wave": best[1], "xcd_affinity": best[2], "us_per_iteration": best[0]}

    # ---------------------------------------------------------------- timing
    def time_step(self, which: str, reps: int = 20) -> float:
        """Mean µs per launch of the column (``"primal"``) or row (``"dual"``) pass at
        the current state: a HIP graph of ``reps`` launches timed with events on
        the engine stream (the state keeps moving; the work per launch does not)."""
        h = C.c_void_p()
        L.check(self.lib.pdlp_family_graph_capture_step(C.byref(self.args), {"primal": 0, "dual": 1}[which], reps,
                                                        self.sh, C.byref(h)), "pdlp_family_graph_capture_step")
        try:
            with torch.cuda.stream(self.stream):
                L.check(self.lib.pdlp_graph_launch(h, self.sh), "warm-up launch")
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(self.stream)
                L.check(self.lib.pdlp_graph_launch(h, self.sh), "timed launch")
                e1.record(self.stream)
            e1.synchronize()
            return e0.elapsed_time(e1) * 1e3 / reps
        finally:
            self.lib.pdlp_graph_destroy(h)
