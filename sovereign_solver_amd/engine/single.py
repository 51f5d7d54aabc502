"""Single-instance PDHG engine: one LP resident in HBM, iterated by HIP kernels.

The host enqueues blocks of ``check_every`` iterations (each block = the
primal and dual kernels ``check_every`` times, then the FP64 KKT check with
the device-side termination/restart decision), captured in HIP graphs of
1, 2, 4, … blocks.  It reads back only the control block (~128 bytes) after
each graph, one graph behind, so the GPU always has the next graph queued.
After convergence the kernels return immediately, so the at most one graph
of surplus work costs little and the reported iteration count is exact.
"""

from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np
import torch

from . import _lib as L
from . import sell
from .device import csr_transpose, dev_tensor, ptr, require_gpu, stream_handle, struct_from_bytes, upload
from ..config import WARMSTART_FAILOVER_ITERS
from .forms import HostForm
from .tune import Tune, heuristic_tune

_STATUS = {L.PDLP_RUNNING: "iteration_limit", L.PDLP_OPTIMAL: "optimal", L.PDLP_NUMERICAL: "numerical_error",
           L.PDLP_INFEASIBLE: "infeasible", L.PDLP_UNBOUNDED: "unbounded"}

# Largest graph (in check blocks).  Small graphs bound the surplus of no-op
# launches after convergence (at most two graphs are in flight).
MAX_GRAPH_BLOCKS = 4


class SingleEngine:
    """Device copy of one :class:`HostForm` plus its PDHG state."""

    def __init__(self, form: HostForm, precision: str = "fp32", tune: Optional[Tune] = None,
                 check_every: int = 50, exact: bool = False, use_graphs: bool = True,
                 scheme: Optional[int] = None):
        if precision not in ("fp32", "fp64"):
            raise ValueError("precision must be 'fp32' or 'fp64'")
        self.lib = L.lib()
        self.device = require_gpu()
        self.form = form
        self.precision = precision
        self.m, self.n, self.nnz = form.m, form.n, form.nnz
        self.tune = tune or heuristic_tune(self.m, self.n, self.nnz, check_every, exact)
        if scheme is not None:
            self.tune = self.tune.with_(scheme=int(scheme))
        if exact and (self.tune.tpr_primal != 1 or self.tune.tpr_dual != 1 or not self.tune.exact):
            raise ValueError("exact mode needs Tune(tpr_primal=1, tpr_dual=1, exact=1)")
        self.use_graphs = use_graphs
        self.stream = torch.cuda.Stream(device=self.device)
        T = torch.float32 if precision == "fp32" else torch.float64
        self.dtype = T
        dv = self.device
        f64 = torch.float64
        i32 = torch.int32
        with torch.cuda.stream(self.stream):
            K = form.K
            # the LP's arrays through one pinned staging buffer, converted on the way
            # (the scale vectors D_r, D_c of a scaled form serve the original LP's KKT)
            (self.row_ptr, self.col_idx, self.val, self.c, self.lo, self.hi, self.b, self.row_kind, self.c64,
             self.b64, self.b_shift, self.lo_shift, self.ub_span, self.row_scale, self.col_scale) = upload(
                [(K.indptr, i32), (K.indices, i32), (K.data, T), (form.c, T), (form.lo, T), (form.hi, T),
                 (form.b, T), (form.row_kind, torch.int8), (form.c, f64), (form.b, f64), (form.b_shift, f64),
                 (form.lo_shift, f64), (form.ub_span, f64), (form.row_scale, f64), (form.col_scale, f64)],
                dv, self.stream)
            # K^T for the primal step, built on the device from the uploaded K
            self.t_row_ptr, self.t_col_idx, self.t_val = csr_transpose(self.row_ptr, self.col_idx, self.val,
                                                                       self.n, self.stream)
            self.x = torch.zeros(self.n, dtype=T, device=dv)
            self.xbar = torch.zeros(self.n, dtype=T, device=dv)
            self.y = torch.zeros(self.m, dtype=T, device=dv)
            self.x_sum = torch.zeros(self.n, dtype=f64, device=dv)
            self.y_sum = torch.zeros(self.m, dtype=f64, device=dv)
            # FP32 engine: FP32 sums over each check block, folded into the FP64 sums at every check
            fp32 = precision == "fp32"
            self.x_psum = torch.zeros(self.n, dtype=torch.float32, device=dv) if fp32 else None
            self.y_psum = torch.zeros(self.m, dtype=torch.float32, device=dv) if fp32 else None
            self.partials = torch.zeros(L.PDLP_PARTIALS_LEN, dtype=f64, device=dv)
            self.ctrl = torch.zeros(C.sizeof(L.PdlpCtrl), dtype=torch.uint8, device=dv)
            # last restart point (adaptive scheme)
            self.x_last = torch.zeros(self.n, dtype=T, device=dv)
            self.y_last = torch.zeros(self.m, dtype=T, device=dv)
            # SELL-64 slices of the rows whose gathers coalesce (not in exact mode: the
            # reference trajectory needs the CSR summation of one lane per row)
            sl = None if exact else sell.build(K)
            self.sell_slices = 0 if sl is None else sl.slices
            if self.sell_slices:
                # column runs (col_k = col_0 + k step) of the slices and the CSR rows: the one-launch
                # dual step then reads one index per row
                (self.sell_ptr, self.sell_rows, self.sell_col, self.sell_val, self.csr_rows,
                 self.sell_stride) = upload(
                    [(sl.ptr, i32), (sl.rows, i32), (sl.col, i32), (sl.val, T),
                     (sl.csr_rows if sl.csr_rows.size else None, i32), (sl.stride, i32)], dv, self.stream)
                # the CSR rows' column steps, found on the device (pdlp_row_strides)
                self.csr_stride = None
                if sl.csr_rows.size:
                    self.csr_stride = torch.empty(sl.csr_rows.size, dtype=i32, device=dv)
                    L.check(self.lib.pdlp_row_strides(int(sl.csr_rows.size), ptr(self.csr_rows), ptr(self.row_ptr),
                                                      ptr(self.col_idx), ptr(self.csr_stride),
                                                      stream_handle(self.stream)), "pdlp_row_strides")
                self._sell_widths = np.diff(sl.ptr) // sell.SLICE
                self._sell_stride_host = sl.stride
                self.n_csr_rows = int(sl.csr_rows.size)
                widths = np.diff(sl.ptr) // sell.SLICE
                self.sell_max_width = int(widths.max())
                # partial sums of the split step: at most ceil(width / 8) chunks per slice
                self.sell_part = torch.zeros(sell.SLICE * self.sell_slices * ((self.sell_max_width + 7) // 8),
                                             dtype=T, device=dv)
                aK = abs(K)
                rs = np.asarray(aK.sum(axis=1)).ravel()
                self.sell_asum = dev_tensor(np.where(sl.rows >= 0, rs[np.maximum(sl.rows, 0)], 1.0), T, dv)
            else:
                self.sell_ptr = self.sell_rows = self.sell_col = self.sell_val = self.csr_rows = None
                self.sell_part = self.sell_asum = self.sell_stride = self.csr_stride = None
                self.n_csr_rows = 0
                self.sell_max_width = 0
            # SELL-64 copy of K^T for the primal step (columns in natural order), when the
            # columns of each slice have nearly equal lengths (padding at most 1/8 of K)
            self.t_sell_slices = 0
            self.t_sell_ptr = self.t_sell_idx = self.t_sell_val = None
            tp = None if exact else sell.build_columns_device(self.t_row_ptr, self.nnz, self.stream)
            if tp is not None:
                self.t_sell_slices = len(tp) - 1
                self.t_sell_ptr = dev_tensor(tp, i32, dv)
                self.t_sell_idx = torch.zeros(int(tp[-1]), dtype=i32, device=dv)
                self.t_sell_val = torch.zeros(int(tp[-1]), dtype=T, device=dv)
                L.check(self.lib.pdlp_csr_to_sell64(L.PDLP_F32 if precision == "fp32" else L.PDLP_F64, self.n,
                                                    ptr(self.t_row_ptr), ptr(self.t_col_idx), ptr(self.t_val),
                                                    ptr(self.t_sell_ptr), ptr(self.t_sell_idx), ptr(self.t_sell_val),
                                                    stream_handle(self.stream)), "pdlp_csr_to_sell64")
        # every entry of K one integer and no padding inside a row of the SELL copies: the SELL
        # steps read no values (k_const)
        self.k_const = 0
        if not exact and K.nnz and (self.sell_slices or self.t_sell_slices):
            v = float(K.data[0])
            ok = v != 0.0 and v == round(v) and abs(v) < 2**24 and bool(np.all(K.data == v))
            if ok and self.sell_slices:
                ok = not sl.padded
            if ok and self.t_sell_slices:
                w_last = int(tp[-1] - tp[-2]) // sell.SLICE
                ok = int(tp[-1]) - (sell.SLICE * self.t_sell_slices - self.n) * w_last == self.nnz
            self.k_const = int(v) if ok else 0
        self.ctrl_host = torch.zeros((2, C.sizeof(L.PdlpCtrl)), dtype=torch.uint8).pin_memory()
        self.stream.synchronize()
        self.dev = L.PdlpDev(
            m=self.m, n=self.n, nnz=self.nnz, prec=L.PDLP_F32 if precision == "fp32" else L.PDLP_F64,
            row_ptr=ptr(self.row_ptr), col_idx=ptr(self.col_idx), val=ptr(self.val),
            t_row_ptr=ptr(self.t_row_ptr), t_col_idx=ptr(self.t_col_idx), t_val=ptr(self.t_val),
            c=ptr(self.c), lo=ptr(self.lo), hi=ptr(self.hi), b=ptr(self.b), row_kind=ptr(self.row_kind),
            c64=ptr(self.c64), b64=ptr(self.b64), b_shift=ptr(self.b_shift), lo_shift=ptr(self.lo_shift),
            ub_span=ptr(self.ub_span), x=ptr(self.x), xbar=ptr(self.xbar), y=ptr(self.y),
            x_sum=ptr(self.x_sum), y_sum=ptr(self.y_sum), x_psum=ptr(self.x_psum), y_psum=ptr(self.y_psum),
            partials=ptr(self.partials), ctrl=ptr(self.ctrl), x_last=ptr(self.x_last), y_last=ptr(self.y_last),
            sell_slices=self.sell_slices, n_csr_rows=self.n_csr_rows, sell_ptr=ptr(self.sell_ptr),
            sell_rows=ptr(self.sell_rows), sell_col=ptr(self.sell_col), sell_val=ptr(self.sell_val),
            csr_rows=ptr(self.csr_rows), sell_max_width=self.sell_max_width, sell_part=ptr(self.sell_part),
            sell_asum=ptr(self.sell_asum), row_scale=ptr(self.row_scale), col_scale=ptr(self.col_scale),
            t_sell_slices=self.t_sell_slices, t_sell_ptr=ptr(self.t_sell_ptr), t_sell_idx=ptr(self.t_sell_idx),
            t_sell_val=ptr(self.t_sell_val), k_const=self.k_const, sell_stride=ptr(self.sell_stride),
            csr_stride=ptr(self.csr_stride))
        self._graphs = {}
        self._graph_tune = None

    # ------------------------------------------------------------------ utils
    @property
    def sh(self):
        return stream_handle(self.stream)

    def _ctune(self, t: Optional[Tune] = None):
        return (t or self.tune).to_c()

    @property
    def dual_index_loads(self) -> int:
        """Column indices one launch of the one-launch SELL dual step reads: every entry's,
        except that a run row (PdlpDev.sell_stride / csr_stride) reads only its first."""
        if not self.sell_slices:
            return int(self.nnz)
        n = int(np.sum(np.where(self._sell_stride_host != 0, sell.SLICE, sell.SLICE * self._sell_widths)))
        if self.csr_stride is not None:
            rl = (self.row_ptr[1:] - self.row_ptr[:-1])[self.csr_rows.long()]
            with torch.cuda.stream(self.stream):
                n += int(torch.where(self.csr_stride != 0, torch.clamp(rl, max=1), rl).sum().item())
        return n

    def set_tune(self, t: Tune) -> None:
        self.tune = t
        self.drop_graphs()

    def drop_graphs(self) -> None:
        for g in self._graphs.values():
            self.lib.pdlp_graph_destroy(g)
        self._graphs = {}

    def __del__(self):
        try:
            self.drop_graphs()
        except Exception:
            pass

    def _graph(self, blocks: int):
        g = self._graphs.get(blocks)
        if g is None:
            h = C.c_void_p()
            ct = self._ctune()
            L.check(self.lib.pdlp_graph_capture(C.byref(self.dev), C.byref(ct), blocks, self.sh, C.byref(h)),
                    "pdlp_graph_capture")
            g = self._graphs[blocks] = h
        return g

    def read_ctrl(self) -> L.PdlpCtrl:
        with torch.cuda.stream(self.stream):
            self.ctrl_host[0].copy_(self.ctrl, non_blocking=True)
        self.stream.synchronize()
        return struct_from_bytes(L.PdlpCtrl, self.ctrl_host[0])

    def write_ctrl(self, st: L.PdlpCtrl) -> None:
        raw = torch.frombuffer(bytearray(bytes(st)), dtype=torch.uint8)
        with torch.cuda.stream(self.stream):
            self.ctrl.copy_(raw.to(self.device, non_blocking=False))

    # ------------------------------------------------------------------ solve
    def start(self, tol: float, x0=None, y0=None, tau_scale: float = 1.0, sigma_scale: float = 1.0) -> None:
        """State for a solve from (x0, y0) (None: the cold start); non-finite entries of a
        start count as the cold start's (x = clip(0, lo, hi), y = 0)."""
        if x0 is not None:
            x0 = np.asarray(x0, float)
            if not np.all(np.isfinite(x0)):
                cold = np.zeros(self.n) if self.lo is None else np.clip(0.0, self.form.lo, self.form.hi)
                x0 = np.where(np.isfinite(x0), x0, cold)
        if y0 is not None:
            y0 = np.asarray(y0, float)
            y0 = np.where(np.isfinite(y0), y0, 0.0)
        with torch.cuda.stream(self.stream):
            if x0 is None and self.lo is None:
                self.x.zero_()
            elif x0 is None:        # cold start: the origin projected onto the box (x = lb when lb >= 0)
                x0 = np.clip(0.0, self.form.lo, self.form.hi)
                self.x.copy_(torch.as_tensor(x0, dtype=self.dtype).to(self.device))
            else:
                self.x.copy_(torch.as_tensor(np.asarray(x0), dtype=self.dtype).to(self.device))
            if y0 is None:
                self.y.zero_()
            else:
                self.y.copy_(torch.as_tensor(np.asarray(y0), dtype=self.dtype).to(self.device))
        st = L.PdlpCtrl()
        st.tol = tol
        st.norm_b = self.form.norm_b
        st.norm_c = self.form.norm_c
        st.tau_scale = tau_scale
        st.sigma_scale = sigma_scale
        self.write_ctrl(st)
        ct = self._ctune()
        L.check(self.lib.pdlp_init(C.byref(self.dev), C.byref(ct), self.sh), "pdlp_init")

    def run_iterations(self, max_iters: int) -> L.PdlpCtrl:
        """Enqueue up to ``max_iters`` iterations; stop early once the device leaves RUNNING."""
        ce = self.tune.check_every
        done = 0
        blocks = 1
        inflight = []          # (slot, event)
        slot = 0
        final = None
        ct = self._ctune()
        while True:
            left = (max_iters - done) // ce
            if left == 0:
                break
            g = min(blocks, left)
            with torch.cuda.stream(self.stream):
                if self.use_graphs:
                    if g & (g - 1):                  # not a power of two: largest power below
                        g = 1 << (g.bit_length() - 1)
                    L.check(self.lib.pdlp_graph_launch(self._graph(g), self.sh), "pdlp_graph_launch")
                else:
                    for _ in range(g):
                        L.check(self.lib.pdlp_run(C.byref(self.dev), C.byref(ct), ce, 1, self.sh), "pdlp_run")
                self.ctrl_host[slot].copy_(self.ctrl, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(self.stream)
            done += g * ce
            inflight.append((slot, ev))
            slot ^= 1
            blocks = min(blocks * 2, MAX_GRAPH_BLOCKS)
            if len(inflight) == 2:
                s0, e0 = inflight.pop(0)
                e0.synchronize()
                st = struct_from_bytes(L.PdlpCtrl, self.ctrl_host[s0])
                if st.status != L.PDLP_RUNNING:
                    final = st
                    break
        if final is None:
            tail = max_iters - done
            if tail > 0:
                with torch.cuda.stream(self.stream):
                    L.check(self.lib.pdlp_run(C.byref(self.dev), C.byref(ct), tail, 0, self.sh), "pdlp_run")
        self.stream.synchronize()
        return self.read_ctrl()

    def initial_weight(self) -> float:
        """Starting primal weight of the adaptive scheme: ‖√σ̂∘b‖ / ‖√τ̂∘c‖.

        The ratio of the right-hand side and objective norms in the metric of the
        diagonal step sizes (τ̂_j = 0.999/colsum_j, σ̂_i = 0.999/rowsum_i); 1 when
        either norm vanishes."""
        f = self.form
        absK = abs(f.K)
        colsum = np.maximum(np.asarray(absK.sum(axis=0)).ravel(), 1e-4)
        rowsum = np.maximum(np.asarray(absK.sum(axis=1)).ravel(), 1e-4)
        nb = float(np.sqrt(np.sum(f.b * f.b * 0.999 / rowsum)))
        nc = float(np.sqrt(np.sum(f.c * f.c * 0.999 / colsum)))
        if nb <= 1e-12 or nc <= 1e-12:
            return 1.0
        return float(np.clip(nb / nc, 1e-6, 1e6))

    def solution(self, st: L.PdlpCtrl):
        """Host copies of (x, y): the FP64 window average when the solve ended on
        the average, else the iterate (the current point, or the last iterate at
        the iteration limit)."""
        with torch.cuda.stream(self.stream):
            if st.status == L.PDLP_OPTIMAL and st.cand == 1:
                w = float(st.apply_window)
                x = (self.x_sum / w).cpu()
                y = (self.y_sum / w).cpu()
            else:
                x = self.x.to(torch.float64).cpu()
                y = self.y.to(torch.float64).cpu()
        self.stream.synchronize()
        return x.numpy(), y.numpy()

    def ray(self, st: L.PdlpCtrl) -> np.ndarray:
        """The ray estimate of an unbounded solve, d = x - (window average), in the
        engine's (scaled) space: the kernels left the current iterate in x and the
        window's FP64 sums in x_sum."""
        with torch.cuda.stream(self.stream):
            d = (self.x.to(torch.float64) - self.x_sum / float(st.apply_window)).cpu()
        self.stream.synchronize()
        return d.numpy()

    def _result(self, st: L.PdlpCtrl) -> dict:
        x, y = self.solution(st)
        iters = (st.done_iter if st.status != L.PDLP_RUNNING else st.iter) + getattr(self, "_offset", 0)
        ray = self.ray(st) if st.status == L.PDLP_UNBOUNDED else None
        return dict(status=_STATUS[st.status], x=x, y=y, iterations=int(iters), rel_kkt=st.last_err,
                    rp=st.rp, rd=st.rd, gap=st.gap, restarts=int(st.restarts), ctrl=st, ray=ray)

    def solve(self, tol: float, max_iters: int, x0=None, y0=None, step_scale: Optional[float] = None,
              failover_iters: int = WARMSTART_FAILOVER_ITERS):
        """One solve.  ``step_scale`` is the starting primal weight w (tau = w tau_hat,
        sigma = sigma_hat / w); None = 1 for the reference scheme, the norm ratio of
        :meth:`initial_weight` for the adaptive one.

        Warm-start failover (``failover_iters`` > 0, the reference's
        WARMSTART_FAILOVER_ITERS): a start given by ``x0``, ``y0`` or ``step_scale`` is
        kept only when its KKT is no worse than the cold start's (when it sets y), and a
        solve whose KKT
        has not fallen below the warm start's by the first check at or after
        ``failover_iters`` iterations, and is at least a tenth of the cold start's, starts
        again from the cold start; the iterations spent before count in the result."""
        w_cold = self.initial_weight() if self.tune.scheme != L.PDLP_SCHEME_REFERENCE else 1.0
        w = w_cold if step_scale is None or not np.isfinite(step_scale) or step_scale <= 0 else float(step_scale)
        self.max_iters = int(max_iters)
        self._offset = 0
        judge = failover_iters > 0 and (x0 is not None or y0 is not None or step_scale is not None)
        if not judge:
            self.start(tol, x0, y0, w, 1.0 / w)
            return self._result(self.run_iterations(max_iters))
        self.start(tol, None, None, w_cold, 1.0 / w_cold)
        cold_err = self.read_ctrl().ref_err
        self.start(tol, x0, y0, w, 1.0 / w)
        st0 = self.read_ctrl()
        start_err = st0.ref_err
        # a start that sets y competes with the cold point at once (with y = 0 the KKT's gap term
        # is near 1 for any x, so x-only starts are judged by their progress only); a start whose
        # KKT is not finite never runs
        if st0.status != L.PDLP_RUNNING or (y0 is not None and not start_err <= cold_err):
            self.start(tol, None, None, w_cold, 1.0 / w_cold)
            return self._result(self.run_iterations(max_iters))
        ce = self.tune.check_every
        first = min(int(max_iters), -(-int(failover_iters) // ce) * ce)
        st = self.run_iterations(first)
        if st.status == L.PDLP_RUNNING and first < max_iters:
            # no progress from the warm start and still far out (one stalled near the tolerance is
            # kept): fail over
            if not st.last_err < start_err and not st.last_err < 0.1 * cold_err:
                self._offset = int(st.iter)
                self.start(tol, None, None, w_cold, 1.0 / w_cold)
                left = int(max_iters) - self._offset
            else:                                        # keep going from where the run stands
                left = int(max_iters) - int(st.iter)
            st = self.run_iterations(left)
        return self._result(st)

    def resume_without_ray_test(self) -> dict:
        """Continue a solve that stopped ``unbounded`` as if its ray had not passed:
        status back to running, the unboundedness test off for the rest of the run, and
        iterations up to the solve's ``max_iters``.  The kernels froze the state when the
        status left running (iterate, window sums, restart point), so the run picks up
        exactly where it stopped."""
        st = self.read_ctrl()
        if st.status != L.PDLP_UNBOUNDED:
            raise RuntimeError("resume_without_ray_test: the last solve did not stop unbounded")
        st.status = L.PDLP_RUNNING
        st.done_iter = 0
        self.write_ctrl(st)
        self.set_tune(self.tune.with_(no_ray=1))
        left = self.max_iters - getattr(self, "_offset", 0) - int(st.iter)
        st = self.run_iterations(left) if left > 0 else self.read_ctrl()
        return self._result(st)

    # --------------------------------------------------------------- timing
    def time_kernel(self, which: str, t: Tune, reps: int = 20) -> float:
        """Mean µs per launch of one step kernel under tuning ``t``: a HIP graph of
        ``reps`` back-to-back launches timed with HIP events on this stream."""
        ct = t.to_c()
        st = L.PdlpCtrl()
        st.tol = 0.0
        st.norm_b, st.norm_c, st.tau_scale, st.sigma_scale = self.form.norm_b, self.form.norm_c, 1.0, 1.0
        self.write_ctrl(st)
        h = C.c_void_p()
        L.check(self.lib.pdlp_graph_capture_step(C.byref(self.dev), C.byref(ct), {"primal": 0, "dual": 1}[which],
                                                 reps, self.sh, C.byref(h)), "pdlp_graph_capture_step")
        try:
            with torch.cuda.stream(self.stream):
                L.check(self.lib.pdlp_graph_launch(h, self.sh), "warm-up launch")
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(self.stream)
                L.check(self.lib.pdlp_graph_launch(h, self.sh), "timed launch")
                e1.record(self.stream)
            self.stream.synchronize()
        finally:
            self.lib.pdlp_graph_destroy(h)
        return e0.elapsed_time(e1) * 1000.0 / reps
