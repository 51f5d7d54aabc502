"""Batched PDHG engine: many independent LPs, persistent HIP launches.

The LPs are packed into one arena of device arrays (each LP's K and K^T by
rows with 16-bit indices, per-column and per-row vectors, FP64 yardstick
data) with a descriptor per LP; ``pdlp_batch_solve`` launches two persistent
kernels, one for the LPs with FP32 iterates and one for the small LPs with
FP64 iterates, whose workgroups take one LP at a time from a work queue
ordered largest-first, so the biggest LPs start first and the small ones fill
in behind them.
"""

from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib as L
from . import scaling as S
from ._lib import PdlpBatch
from .device import dev_tensor, ptr, require_gpu, stream_handle
from .forms import ROW_EQ, ROW_LE
from ..problem import Problem

LP_DTYPE = np.dtype([
    ("m", np.int32), ("n", np.int32), ("tpr", np.int32), ("flags", np.int32), ("tpc", np.int32),
    ("reserved", np.int32),
    ("row_off", np.int64), ("col_off", np.int64), ("rp_off", np.int64), ("nnz_off", np.int64), ("cp_off", np.int64),
    ("norm_b", np.float64), ("norm_c", np.float64), ("tau_scale", np.float64), ("sigma_scale", np.float64),
], align=True)

OUT_DTYPE = np.dtype([
    ("status", np.int32), ("iterations", np.int32), ("restarts", np.int32), ("checks", np.int32),
    ("rel_kkt", np.float64), ("rp", np.float64), ("rd", np.float64), ("gap", np.float64),
    ("p_obj", np.float64), ("d_obj", np.float64), ("weight", np.float64),
], align=True)


_STATUS = {L.PDLP_RUNNING: "iteration_limit", L.PDLP_OPTIMAL: "optimal", L.PDLP_NUMERICAL: "numerical_error",
           L.PDLP_INFEASIBLE: "infeasible", L.PDLP_SKIPPED: "skipped"}


ENTRIES_PER_LANE = 8     # entries each lane loads at once in a pass (EL in pdlp_batch.hip)


def lanes_per_row(mean_len: float) -> int:
    """Lanes per line (row or column) for a mean line length: the smallest power
    of two in [1, 64] whose lanes hold a line of mean + 2 sqrt(mean) entries
    (most lines of a random sparse matrix) in one load of eight entries each.
    Fewer lanes per line mean more lines per round."""
    want = (mean_len + 2.0 * np.sqrt(max(mean_len, 0.0))) / ENTRIES_PER_LANE
    t = 1
    while t < want and t < 64:
        t *= 2
    return t


def lds_bytes_of(m, n):
    """LDS bytes one LP needs in the persistent kernel with FP32 iterates: n
    doubles (xbar / FP64 primal point), m floats (y), m+1 and n+1 line pointers."""
    return 12 * np.asarray(n) + 8 * np.asarray(m) + 8


def lds_bytes_x_in_lds(m, n):
    """LDS bytes with FP32 iterates and x, its partial sum, tau_hat and c in LDS
    as well (the FP32 part when every one of its LPs fits)."""
    return 28 * np.asarray(n) + 8 * np.asarray(m) + 8


def lds_bytes_resident(m, n):
    """LDS bytes with FP64 iterates and x in LDS (the FP64 part): x, xbar (n
    doubles each), y (m doubles) and the line pointers."""
    return 20 * np.asarray(n) + 12 * np.asarray(m) + 8


# LPs whose FP64 part needs at most this much LDS (m up to about 260 when
# n = 2.5m) run with FP64 iterates; larger ones with FP32 iterates on 512
# threads, which finish a mid-size LP (m ~ 1,000) 1.6x sooner than the FP64
# part's 256 threads.  On the uneven 4,096-LP batch: 0.600 s at 16 KB,
# 0.606 s at 24 KB, 0.614 s at 40 KB, 0.711 s at 64 KB (parts back to back,
# earlier kernels); with the parts concurrent, 0.481 s at 16 KB, 0.482 s at
# 24 KB, 0.484 s at 32 KB, 0.494 s at 40 KB.
RESIDENT_LDS = 16 * 1024


def _segment_sums(v: np.ndarray, starts: np.ndarray, lens: np.ndarray) -> np.ndarray:
    """Sums of the segments v[starts[k] : starts[k] + lens[k]] (0 for empty ones).
    The segments tile v in order (each starts where the previous one ends)."""
    out = np.zeros(len(starts))
    nz = lens > 0
    if nz.any():
        st = starts[nz]
        sums = np.add.reduceat(v, st)          # segment i runs to the next non-empty start
        end = int(st[-1] + lens[nz][-1])
        sums[-1] -= v[end:].sum()              # the last one runs to the end of v
        out[nz] = sums
    return out


class PackedBatch:
    """Host-side arena of B box-form LPs (same layout as the device arena).

    ``scaling`` (a method of :mod:`.scaling`): each stored matrix, and with it
    its LPs' b, c and bounds, is scaled (K = D_r K0 D_c); ``row_scale`` and
    ``col_scale`` hold D_r and D_c per row and column of the arena, the
    normalisers stay the original LPs', and the engine maps starting points
    and solutions between the two spaces."""

    def __init__(self, problems: Sequence[Problem], scaling: str = "none", transpose: str = "device"):
        if scaling not in S.METHODS:
            raise ValueError(f"scaling must be one of {S.METHODS}, got {scaling!r}")
        if transpose not in ("device", "host"):
            raise ValueError("transpose must be 'device' or 'host'")
        self.problems = list(problems)
        self.scaling = scaling
        B = len(self.problems)
        # LPs whose (A_ub, A_eq) are the same objects share one stored CSR
        mat_of = np.empty(B, np.int64)
        uniq = {}
        stored = []
        for k, p in enumerate(self.problems):
            key = (id(p.A_ub), id(p.A_eq))
            if key not in uniq:
                K = p.A_eq if p.n_ineq == 0 else (p.A_ub if p.n_eq == 0 else
                                                  sp.vstack([p.A_ub, p.A_eq], format="csr"))
                if not (sp.isspmatrix_csr(K) and K.dtype == np.float64):
                    K = sp.csr_matrix(K, dtype=np.float64)
                if not K.has_canonical_format:       # sums duplicates and sorts each row's indices
                    K = K.copy()
                    K.sum_duplicates()
                uniq[key] = len(stored)
                stored.append(K)
            mat_of[k] = uniq[key]
        s_m = np.array([K.shape[0] for K in stored], np.int64)
        s_n = np.array([K.shape[1] for K in stored], np.int64)
        s_z = np.array([K.nnz for K in stored], np.int64)
        ms, ns, nnzs = s_m[mat_of], s_n[mat_of], s_z[mat_of]
        bad = np.flatnonzero((ms == 0) | (ns == 0))
        if bad.size:
            raise ValueError(f"LP {int(bad[0])}: the batched engine needs m > 0 and n > 0")
        big = np.flatnonzero(lds_bytes_of(ms, ns) > 160 * 1024)      # also keeps m, n below 2^15
        if big.size:
            raise ValueError(f"LP {int(big[0])} (m={int(ms[big[0]])}, n={int(ns[big[0]])}) needs more than a CU's "
                             f"160 KB of LDS; solve it with the family or single-instance engine")
        self.ms, self.ns, self.nnzs = ms, ns, nnzs
        self.n_matrices = len(stored)
        row_off = np.concatenate([[0], np.cumsum(ms)[:-1]])
        col_off = np.concatenate([[0], np.cumsum(ns)[:-1]])
        s_rp = np.concatenate([[0], np.cumsum(s_m + 1)[:-1]])
        s_nz = np.concatenate([[0], np.cumsum(s_z)[:-1]])
        rp_off = s_rp[mat_of]
        nnz_off = s_nz[mat_of]
        self.row_off, self.col_off, self.nnz_off = row_off, col_off, nnz_off
        M, N, Z = int(ms.sum()), int(ns.sum()), int(nnzs.sum())
        if int(s_z.sum()) >= 2**31 or N >= 2**31:
            raise ValueError("batch beyond int32 indexing of the arena")
        self.M, self.N, self.Z = M, N, Z
        self.Z_stored = int(s_z.sum())
        # scaled copies of the stored matrices (same pattern); b_shift below uses the originals
        if scaling != "none":
            sc = [S.scales(K, scaling) for K in stored]
            stored_orig, stored = stored, [S.scale_matrix(K, dr, dc) for K, (dr, dc) in zip(stored, sc)]
        else:
            stored_orig = stored
        self.row_ptr = np.concatenate([K.indptr for K in stored]).astype(np.int32)
        col_local = np.concatenate([K.indices for K in stored])
        self.col_idx = col_local.astype(np.uint16)
        self.val = np.concatenate([K.data for K in stored])
        s_ro = np.concatenate([[0], np.cumsum(s_m)[:-1]])        # first row / column of each matrix
        s_co = np.concatenate([[0], np.cumsum(s_n)[:-1]])
        s_cp = np.concatenate([[0], np.cumsum(s_n + 1)[:-1]])
        self.cp_off = s_cp[mat_of]
        # per stored matrix (mats + 1 running offsets): rows, columns and entries in the
        # block-diagonal union, row pointers and column pointers in the arena
        cum = lambda v: np.concatenate([[0], np.cumsum(v)]).astype(np.int32)
        self.arena = {"row_off": cum(s_m), "col_off": cum(s_n), "nz_off": cum(s_z), "rp_off": cum(s_m + 1),
                      "cp_off": cum(s_n + 1)}
        if transpose == "device":
            # K^T built by the engine on the device (pdlp_arena_transpose), bitwise the host's below
            self.t_row_idx = self.t_val = self.t_col_ptr = None
        else:
            # K^T by rows (K by columns) of every stored matrix at once: one transpose of their
            # block-diagonal union (SciPy's counting sort; each column's entries stay in row order)
            gptr = np.concatenate([np.concatenate([K.indptr[:-1] + z for K, z in zip(stored, s_nz)]),
                                   [int(s_z.sum())]]).astype(np.int32)
            gcol = col_local.astype(np.int32)
            gcol += np.repeat(s_co.astype(np.int32), s_z)
            BT = sp.csr_matrix((self.val, gcol, gptr), shape=(int(s_m.sum()), int(s_n.sum()))).T.tocsr()
            col_mat = np.repeat(np.arange(len(stored)), s_n)            # the matrix of each global column
            rows = BT.indices.astype(np.int32, copy=False)
            rows -= np.repeat(s_ro[col_mat].astype(np.int32), np.diff(BT.indptr))
            self.t_row_idx = rows.astype(np.uint16)
            self.t_val = BT.data
            # per matrix its n + 1 column pointers, relative to its first entry
            first = np.repeat(s_co, s_n + 1) + np.concatenate([np.arange(n + 1) for n in s_n])
            self.t_col_ptr = (BT.indptr[first] - np.repeat(BT.indptr[s_co], s_n + 1)).astype(np.int32)

        P = self.problems
        c = np.concatenate([np.asarray(p.c, float) for p in P])
        lo = np.concatenate([np.asarray(p.lb, float) for p in P])
        hi = np.concatenate([np.asarray(p.ub, float) for p in P])
        b = np.concatenate([np.concatenate([np.asarray(p.b_ub, float), np.asarray(p.b_eq, float)]) for p in P])
        m_ub = np.array([p.n_ineq for p in P], np.int64)
        lp_of_row = np.repeat(np.arange(B), ms)
        kind = np.where(np.arange(M) - row_off[lp_of_row] < m_ub[lp_of_row], ROW_LE, ROW_EQ).astype(np.int8)
        lbf = np.where(np.isfinite(lo), lo, 0.0)
        b_shift = b.copy()
        for k in np.flatnonzero(_segment_sums(np.abs(lbf), col_off, ns) > 0):   # only LPs with a lower shift
            K = stored_orig[mat_of[k]]
            b_shift[row_off[k]: row_off[k] + ms[k]] -= K @ lbf[col_off[k]: col_off[k] + ns[k]]
        fin = np.isfinite(hi)
        span2 = np.where(fin, (hi - lbf) ** 2, 0.0)
        desc = np.zeros(B, LP_DTYPE)
        desc["m"], desc["n"] = ms, ns
        desc["tpr"] = [lanes_per_row(z / m) for z, m in zip(nnzs, ms)]
        desc["tpc"] = [lanes_per_row(z / n) for z, n in zip(nnzs, ns)]
        desc["row_off"], desc["col_off"], desc["rp_off"], desc["nnz_off"] = row_off, col_off, rp_off, nnz_off
        desc["cp_off"] = self.cp_off
        desc["norm_b"] = np.maximum(1.0, np.sqrt(_segment_sums(b_shift * b_shift, row_off, ms) +
                                                 _segment_sums(span2, col_off, ns)))
        desc["norm_c"] = np.maximum(1.0, np.sqrt(_segment_sums(c * c, col_off, ns)))
        desc["tau_scale"], desc["sigma_scale"] = 1.0, 1.0
        lo_shift = lbf
        ub_span = np.where(fin, hi - lo_shift, np.inf)
        self.row_scale = self.col_scale = None
        if scaling != "none":                    # the LPs of the scaled matrices (normalisers above stay)
            rs = np.concatenate([sc[mat_of[k]][0] for k in range(B)])
            cs = np.concatenate([sc[mat_of[k]][1] for k in range(B)])
            c, lo, hi, b = c * cs, lo / cs, hi / cs, b * rs
            b_shift, lo_shift, ub_span = b_shift * rs, lo_shift / cs, ub_span / cs
            self.row_scale, self.col_scale = rs, cs
        self.c, self.lo, self.hi, self.b, self.kind = c, lo, hi, b, kind
        self.lo_shift, self.ub_span, self.b_shift = lo_shift, ub_span, b_shift
        self.desc = desc
        self.set_split(RESIDENT_LDS)

    def set_split(self, resident_lds: int) -> None:
        """The FP32/FP64 split: LPs whose FP64-part layout needs at most ``resident_lds``
        bytes of LDS run with FP64 iterates.  Sets the work order (FP32 part largest
        first, then the FP64 part largest first) and both parts' LDS sizes."""
        ms, ns, nnzs = self.ms, self.ns, self.nnzs
        self.split = int(resident_lds)
        self.resident = lds_bytes_resident(ms, ns) <= self.split
        by_size = np.argsort(-nnzs, kind="stable")
        self.order = np.concatenate([by_size[~self.resident[by_size]], by_size[self.resident[by_size]]]).astype(np.int32)
        self.n_resident = int(self.resident.sum())
        r16 = lambda v: int((int(v) + 15) // 16 * 16)
        # FP32 part: x, its partial sum, tau_hat and c in LDS too when every LP of the part
        # fits (one 512-thread workgroup per CU leaves the CU's whole LDS to it)
        fp32 = ~self.resident
        self.set_x_in_lds(bool(fp32.any() and lds_bytes_x_in_lds(ms, ns)[fp32].max() <= 160 * 1024))
        self.lds_resident = r16(lds_bytes_resident(ms, ns)[self.resident].max()) if self.n_resident else 16
        # the largest need of the FP64 part, checked by pdlp_batch_solve against its LDS
        self.lds_need_resident = int(lds_bytes_resident(ms, ns)[self.resident].max()) if self.n_resident else 0

    def set_x_in_lds(self, on: bool) -> None:
        """FP32 part layout: x, its partial sum, tau_hat and c in LDS (``on``) or in global
        memory; sets the part's LDS per workgroup and its largest need to match."""
        fp32 = ~self.resident
        need = (lds_bytes_x_in_lds if on else lds_bytes_of)(self.ms, self.ns)
        if on and fp32.any() and need[fp32].max() > 160 * 1024:
            raise ValueError("an LP of the FP32 part does not fit in 160 KiB with x in LDS")
        self.x_in_lds = bool(on)
        self.lds_need = int(need[fp32].max()) if fp32.any() else 0
        self.lds_bytes = int((self.lds_need + 15) // 16 * 16) if fp32.any() else 16

    def algorithmic_bytes_per_iteration(self) -> float:
        """Bytes one iteration of every LP moves at minimum: K and K^T once each
        (values and indices; line pointers stay in LDS) and the per-column and
        per-row state."""
        return float(self.Z * 16 + self.N * (4 * 6 + 4 * 2) + self.M * 5)


class BatchEngine:
    """Device arena of a :class:`PackedBatch`.  ``scheme``: "adaptive" (default)
    or "reference" restarts for every LP (see ``include/pdlp_hip.h``)."""

    def __init__(self, packed: PackedBatch, block: int = 0, grid: int = 0, scheme: str = "adaptive",
                 block_resident: int = 256, check_every: int = 50, lines: int = 3):
        if scheme not in ("adaptive", "reference", "adaptive_fixed", "adaptive_start"):
            raise ValueError("scheme must be 'adaptive', 'adaptive_fixed', 'adaptive_start' or 'reference'")
        self.lib = L.lib()
        self.device = require_gpu()
        self.pk = packed
        self.adaptive = scheme != "reference"
        self.flags = {"reference": 0, "adaptive": 1, "adaptive_fixed": 3, "adaptive_start": 5}[scheme]
        # FP32 part (the large LPs): 512 threads, one workgroup per CU, measured
        # faster than two of 256 on the uneven batch; FP64 part: 256 threads
        self.block = block or 512
        self.block_resident = block_resident
        self.grid = grid
        self.check_every = int(check_every)
        self.lines = int(lines)
        self.stream = torch.cuda.Stream(device=self.device)
        dv = self.device
        f32, f64, i32 = torch.float32, torch.float64, torch.int32
        pk = packed
        with torch.cuda.stream(self.stream):
            self.t = {
                "lps": torch.from_numpy(pk.desc.view(np.uint8).copy()).to(dv),
                "order": dev_tensor(pk.order, i32, dv),
                "queue": torch.zeros(2, dtype=i32, device=dv),
                "row_ptr": dev_tensor(pk.row_ptr, i32, dv),
                "col_idx": dev_tensor(pk.col_idx.view(np.int16), torch.int16, dv),
                "val": dev_tensor(pk.val, f32, dv),
                "c": dev_tensor(pk.c, f32, dv), "lo": dev_tensor(pk.lo, f32, dv), "hi": dev_tensor(pk.hi, f32, dv),
                "c64": dev_tensor(pk.c, f64, dv), "lo_shift": dev_tensor(pk.lo_shift, f64, dv),
                "ub_span": dev_tensor(pk.ub_span, f64, dv),
                "b": dev_tensor(pk.b, f32, dv), "row_kind": dev_tensor(pk.kind, torch.int8, dv),
                "b64": dev_tensor(pk.b, f64, dv), "b_shift": dev_tensor(pk.b_shift, f64, dv),
                "x": torch.zeros(pk.N, dtype=f32, device=dv), "x_psum": torch.zeros(pk.N, dtype=f32, device=dv),
                "x_sum": torch.zeros(pk.N, dtype=f64, device=dv), "tau": torch.zeros(pk.N, dtype=f32, device=dv),
                "y": torch.zeros(pk.M, dtype=f32, device=dv), "y_sum": torch.zeros(pk.M, dtype=f64, device=dv),
                "out": torch.zeros(len(pk.problems) * OUT_DTYPE.itemsize, dtype=torch.uint8, device=dv),
                "x_last": torch.zeros(pk.N, dtype=f32, device=dv), "y_last": torch.zeros(pk.M, dtype=f32, device=dv),
                "y_psum": torch.zeros(pk.M, dtype=f32, device=dv),
            }
            if pk.row_scale is not None:
                self.t["row_scale"] = dev_tensor(pk.row_scale, f64, dv)
                self.t["col_scale"] = dev_tensor(pk.col_scale, f64, dv)
            if pk.t_row_idx is not None:
                self.t["t_col_ptr"] = dev_tensor(pk.t_col_ptr, i32, dv)
                self.t["t_row_idx"] = dev_tensor(pk.t_row_idx.view(np.int16), torch.int16, dv)
                self.t["t_val"] = dev_tensor(pk.t_val, f32, dv)
            else:
                self._arena_transpose()
        self.stream.synchronize()

    def _arena_transpose(self) -> None:
        """K^T of every stored matrix on the device (``pdlp_arena_transpose``) from the
        uploaded K: per matrix bitwise the host packer's transpose."""
        pk, dv, t = self.pk, self.device, self.t
        a = {k: dev_tensor(v, torch.int32, dv) for k, v in pk.arena.items()}
        mats = len(pk.arena["row_off"]) - 1
        rows, cols, nnz = int(pk.arena["row_off"][-1]), int(pk.arena["col_off"][-1]), int(pk.arena["nz_off"][-1])
        need = C.c_size_t(0)
        L.check(self.lib.pdlp_arena_transpose_work_bytes(mats, rows, cols, nnz, C.byref(need)),
                "pdlp_arena_transpose_work_bytes")
        work = torch.empty(max(int(need.value), 1), dtype=torch.uint8, device=dv)
        t["t_col_ptr"] = torch.empty(int(pk.arena["cp_off"][-1]), dtype=torch.int32, device=dv)
        t["t_row_idx"] = torch.empty(nnz, dtype=torch.int16, device=dv)
        t["t_val"] = torch.empty(nnz, dtype=torch.float32, device=dv)
        L.check(self.lib.pdlp_arena_transpose(
            L.PDLP_F32, mats, rows, cols, nnz, ptr(a["row_off"]), ptr(a["col_off"]), ptr(a["nz_off"]),
            ptr(a["rp_off"]), ptr(a["cp_off"]), ptr(t["row_ptr"]), ptr(t["col_idx"]), ptr(t["val"]),
            ptr(t["t_col_ptr"]), ptr(t["t_row_idx"]), ptr(t["t_val"]), ptr(work), int(need.value),
            stream_handle(self.stream)), "pdlp_arena_transpose")
        self.stream.synchronize()                # work and the offsets are freed on return

    def configure(self, check_every: Optional[int] = None, block: Optional[int] = None,
                  lines: Optional[int] = None, split: Optional[int] = None,
                  block_resident: Optional[int] = None) -> None:
        """Launch configuration (None keeps a value): check cadence, threads per
        workgroup of the FP32 part (256, 512), its lines per lane group and round (2-4),
        the FP32/FP64 split (:meth:`PackedBatch.set_split`) and the FP64 part's
        threads per workgroup (256, 512).  None of these changes an LP's arithmetic
        except the cadence (when the checks run) and the split (its iterate precision)."""
        if check_every is not None:
            if int(check_every) <= 0:
                raise ValueError("check_every must be positive")
            self.check_every = int(check_every)
        if block is not None:
            if int(block) not in (256, 512):
                raise ValueError("block must be 256 or 512")
            self.block = int(block)
        if lines is not None:
            if int(lines) not in (2, 3, 4):
                raise ValueError("lines must be 2, 3 or 4")
            self.lines = int(lines)
        if block_resident is not None:
            if int(block_resident) not in (256, 512):
                raise ValueError("block_resident must be 256 or 512")
            self.block_resident = int(block_resident)
        if split is not None and int(split) != self.pk.split:
            self.pk.set_split(int(split))
            with torch.cuda.stream(self.stream):
                self.t["order"].copy_(torch.from_numpy(self.pk.order).to(self.device))
            self.stream.synchronize()

    def set_warm_start(self, x0: Optional[np.ndarray] = None, y0: Optional[np.ndarray] = None,
                       step_scale: Optional[np.ndarray] = None) -> None:
        """Packed starting point (length N / M arrays in the LPs' own space; None =
        cold start) and per-LP starting primal weights (None = 1 for the
        reference scheme, the norm-ratio estimate for the adaptive one)."""
        if self.pk.col_scale is not None:        # into the scaled LPs' space
            x0 = None if x0 is None else np.asarray(x0, float) / self.pk.col_scale
            y0 = None if y0 is None else np.asarray(y0, float) / self.pk.row_scale
        with torch.cuda.stream(self.stream):
            if x0 is None:          # cold start: the origin projected onto the box (x = lb when lb >= 0)
                self.t["x"].copy_(torch.from_numpy(np.clip(0.0, self.pk.lo, self.pk.hi).astype(np.float32))
                                  .to(self.device))
            else:
                self.t["x"].copy_(torch.from_numpy(np.asarray(x0, np.float32)).to(self.device))
            if y0 is None:
                self.t["y"].zero_()
            else:
                self.t["y"].copy_(torch.from_numpy(np.asarray(y0, np.float32)).to(self.device))
            desc = self.pk.desc.copy()
            desc["flags"] = self.flags
            # NaN: the kernel computes the norm-ratio starting weight of the adaptive scheme
            w = np.full(len(desc), np.nan) if self.adaptive else np.ones(len(desc))
            if step_scale is not None:            # NaN entries keep the default
                ss = np.asarray(step_scale, float)
                w = np.where(np.isfinite(ss), ss, w)
            desc["tau_scale"] = w
            desc["sigma_scale"] = 1.0 / w
            self.t["lps"].copy_(torch.from_numpy(desc.view(np.uint8).copy()).to(self.device))

    def _args(self, tol: float, max_iters: int, check_every: int) -> PdlpBatch:
        t = self.t
        a = PdlpBatch(count=len(self.pk.problems), check_every=check_every, max_iters=max_iters, block=self.block,
                      lds_bytes=self.pk.lds_bytes, grid=self.grid, n_resident=self.pk.n_resident,
                      block_resident=self.block_resident, lds_resident=self.pk.lds_resident, grid_resident=0,
                      x_in_lds=int(self.pk.x_in_lds), lds_need=self.pk.lds_need,
                      lds_need_resident=self.pk.lds_need_resident, lines=self.lines, tol=tol)
        for name in ("lps", "order", "queue", "row_ptr", "col_idx", "val", "t_col_ptr", "t_row_idx", "t_val",
                     "c", "lo", "hi", "c64", "lo_shift",
                     "ub_span", "b", "row_kind", "b64", "b_shift", "x", "x_psum", "x_sum", "tau", "y", "y_sum",
                     "out", "x_last", "y_last", "y_psum"):
            setattr(a, name, ptr(t[name]))
        if "row_scale" in t:
            a.row_scale, a.col_scale = ptr(t["row_scale"]), ptr(t["col_scale"])
        return a

    def launch(self, tol: float, max_iters: int, check_every: Optional[int] = None, parts: int = 0) -> None:
        """Enqueue one batched solve on the engine stream (inputs already in place).
        ``parts``: 1 = only the FP32 part, 2 = only the FP64 part, 0 = both;
        ``check_every`` None = the engine's cadence."""
        self._a = self._args(tol, max_iters, self.check_every if check_every is None else check_every)
        self._a.parts = parts
        with torch.cuda.stream(self.stream):
            if parts in (0, 3):
                self.t["queue"].zero_()
            else:
                self.t["queue"][parts - 1].zero_()
            L.check(self.lib.pdlp_batch_solve(C.byref(self._a), stream_handle(self.stream)), "pdlp_batch_solve")

    def results(self):
        self.stream.synchronize()
        out = np.frombuffer(self.t["out"].cpu().numpy().tobytes(), dtype=OUT_DTYPE)
        xs = self.t["x_sum"].cpu().numpy()
        ys = self.t["y_sum"].cpu().numpy()
        if self.pk.col_scale is not None:        # back to the LPs' own space
            xs, ys = xs * self.pk.col_scale, ys * self.pk.row_scale
        return out, xs, ys

    def solve(self, tol: float, max_iters: int, check_every: Optional[int] = None, x0=None, y0=None,
              step_scale=None):
        """Cold start (x = y = 0) unless a packed warm start is given."""
        self.set_warm_start(x0, y0, step_scale)
        self.launch(tol, max_iters, check_every)
        return self.results()

    def unpack(self, out, xs, ys):
        """Per-LP (status, x, y, out-record) in each LP's own variable/row space."""
        pk = self.pk
        res = []
        for k in range(len(pk.problems)):
            x = xs[pk.col_off[k]: pk.col_off[k] + pk.ns[k]]
            y = ys[pk.row_off[k]: pk.row_off[k] + pk.ms[k]]
            res.append((_STATUS[int(out["status"][k])], x, y, out[k]))
        return res
