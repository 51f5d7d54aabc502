"""Node model: one MLP applied to every variable of every instance of an LP family.

The instance-level network of :mod:`.warmstart` maps an instance's features to
its whole solution through one linear output layer, so its predictions lie in
a d_h-dimensional affine subspace (plus the ReLU), however well it is trained;
on the 200 x 24 dispatch family the best rank-512 fit of 32,768 PDHG solutions
still leaves 8.7% of x and 25% of y.  The node model instead runs the same
3-layer, width-d_h MLP on one row per *node* (a variable of one instance),
with weights shared by all nodes and instances:

* the family supplies a :class:`NodeSpec`: an index table that builds each
  node's input row from a compact per-instance vector (its data and cheap
  derived quantities) and a shared vector (per-variable constants), and a
  target map giving, per node and output channel, which entry of the solution
  [x*, y*] it predicts and with what scale;
* ``pdlp_node_gather`` (HIP) builds the BF16 input rows and FP32 target rows
  on the device, drawing training rows uniformly from all (instance, node)
  pairs by a counter-based hash;
* the MLP kernels of :mod:`.mlp` train and run the network (MFMA GEMMs with
  M = rows, K-split weight gradients);
* ``pdlp_node_assemble`` (HIP) maps the per-node outputs back to a start
  [x, y, log primal weight] per instance: an entry of [x, y] predicted by
  several nodes (a period's demand dual, by every generator of the period)
  takes their mean; the output normalisation is folded into the map;
* the start goes to the family engine through ``pdlp_family_set_start``
  (projection onto the bounds and dual signs) and PDHG finishes and judges
  every solve, with the warm-start failover of the engine.

A second stage (``prev=`` the first model, ``neighbours=`` a node
neighbourhood from the family) is the same network again whose input row also
holds the first stage's normalised outputs at the node's neighbours (on the
dispatch family: the same generator at periods t-2..t+2), gathered from the
first stage's output as a second per-instance source: one round of message
passing, which lets the ramp couplings between periods correct the first
stage's per-period guesses.

The primal weight is one constant per family, chosen by solving a validation
set of training instances from the network's starts at a few weights
(:meth:`NodeModel.calibrate_weight`): the distance-bound weight
‖x0 − x*‖/‖y0 − y*‖ of the starts measured ~15 on the dispatch family while
the fastest constant was ~2.
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from ..engine import _lib as L
from ..engine.device import stream_handle
from ..lp.batch import results_of
from .mlp import WarmStartMLP
from .warmstart import _engine_for, padded_width


@dataclass
class NodeSpec:
    nodes: int                    # nodes per instance
    table: np.ndarray             # [nodes, width] int32: < src_len per-instance vector, else shared, < 0 zero
    shared: np.ndarray            # [shared_len] float32
    src_len: int                  # per-instance vector length
    target_index: np.ndarray      # [nodes, channels] int32 into [x (n) | y (m)], -1 none
    target_scale: np.ndarray      # [nodes, channels] float32 (non-zero where the index is set)

    @property
    def width(self) -> int:
        return int(self.table.shape[1])

    @property
    def channels(self) -> int:
        return int(self.target_index.shape[1])


def assembly_map(spec: NodeSpec, length: int, mean: np.ndarray, std: np.ndarray):
    """CSR map from normalised node outputs to the entries [0, length) of [x | y]:
    entry i = mean over the cells (node j, channel k) with target_index[j, k] = i of
    (std[k] o[j, k] + mean[k]) / target_scale[j, k].  Returns (indptr, node, chan,
    weight, bias) with weight and bias folded from mean and std."""
    j, k = np.nonzero(spec.target_index >= 0)
    i = spec.target_index[j, k].astype(np.int64)
    if np.any(i >= length):
        raise ValueError("target index beyond the start vector")
    s = spec.target_scale[j, k].astype(np.float64)
    if np.any(s == 0):
        raise ValueError("target_scale must be non-zero where target_index is set")
    cnt = np.bincount(i, minlength=length).astype(np.float64)
    inv = 1.0 / (s * cnt[i])
    w = std[k] * inv
    b = np.bincount(i, weights=mean[k] * inv, minlength=length)
    order = np.argsort(i, kind="stable")
    indptr = np.zeros(length + 1, np.int32)
    indptr[1:] = np.cumsum(np.bincount(i, minlength=length))
    return (indptr, j[order].astype(np.int32), k[order].astype(np.int32), w[order].astype(np.float32),
            b.astype(np.float32))


def _ptr(t: Optional[torch.Tensor]) -> int:
    return 0 if t is None else t.data_ptr()


class NodeModel:
    """The node network of a family with its normalisers, assembly map and primal
    weight; with ``prev`` a second stage reading ``prev``'s outputs at each node's
    ``neighbours`` ([nodes, K] node indices, -1 none)."""

    def __init__(self, spec: NodeSpec, n: int, m: int, d_h: int = 512, seed: int = 0, device=None,
                 prev: Optional["NodeModel"] = None, neighbours: Optional[np.ndarray] = None):
        self.spec, self.n, self.m = spec, n, m
        self.prev = prev
        table = np.asarray(spec.table, np.int32)
        if prev is not None:
            if neighbours is None or neighbours.shape[0] != spec.nodes:
                raise ValueError("a second stage needs neighbours [nodes, K]")
            nb = np.asarray(neighbours, np.int64)
            ch = np.arange(prev.spec.channels)
            idx = nb[:, :, None] * prev.d_out + ch[None, None, :]
            idx = np.where(nb[:, :, None] >= 0, spec.src_len + spec.shared.size + idx, -1)
            table = np.concatenate([table, idx.reshape(spec.nodes, -1).astype(np.int32)], axis=1)
        self.width = int(table.shape[1])
        self.d_in, self.d_out = padded_width(self.width), padded_width(spec.channels)
        self.mlp = WarmStartMLP(self.d_in, d_h, self.d_out, seed=seed, device=device,
                                stream=None if prev is None else prev.mlp.stream)
        self.lib = self.mlp.lib
        self.device = self.mlp.device
        dv = self.device
        with torch.cuda.stream(self.mlp.stream):
            self.table = torch.from_numpy(np.ascontiguousarray(table)).to(dv)
            self.shared = torch.from_numpy(np.ascontiguousarray(spec.shared, np.float32)).to(dv)
            self.t_index = torch.from_numpy(np.ascontiguousarray(spec.target_index, np.int32)).to(dv)
            self.t_scale = torch.from_numpy(np.ascontiguousarray(spec.target_scale, np.float32)).to(dv)
            w = np.zeros(self.d_out, np.float32)
            w[: spec.channels] = 1.0
            self.loss_weight = torch.from_numpy(w).to(dv)
        self.in_mean = self.in_inv_std = self.out_mean = self.out_inv_std = None
        self.out_std_host = self.out_mean_host = None
        self.log_weight: Optional[float] = None
        self._map = None
        self.train_seconds = 0.0
        self.train_gpu_seconds = 0.0
        self.train_flop = 0.0

    @property
    def sh(self):
        return stream_handle(self.mlp.stream)

    # ---------------------------------------------------------------- gathers
    def _gather_inputs(self, V: torch.Tensor, out: torch.Tensor, count: int = 0, seed: int = 0,
                       normalise: bool = True, rows: Optional[torch.Tensor] = None,
                       P: Optional[torch.Tensor] = None) -> None:
        """Input rows from per-instance vectors V [B, src_len] (and, for a second stage,
        the first stage's outputs P [B, nodes * prev.d_out])."""
        sp = self.spec
        if (P is None) != (self.prev is None):
            raise ValueError("P (the first stage's outputs) is given exactly for a second stage")
        if P is not None and (P.shape[0] != V.shape[0] or P.shape[1] != sp.nodes * self.prev.d_out):
            raise ValueError("P must be [B, nodes * prev.d_out]")
        g = L.PdlpNodeGather(batch=V.shape[0], nodes=sp.nodes, width=self.width, ld_out=self.d_in,
                             src_len=sp.src_len, shared_len=self.shared.numel(),
                             src2_len=0 if P is None else P.shape[1], count=count, seed=seed,
                             src=V.data_ptr(), shared=self.shared.data_ptr(), src2=_ptr(P),
                             table=self.table.data_ptr(), scale=0, mean=_ptr(self.in_mean) if normalise else 0,
                             inv_std=_ptr(self.in_inv_std) if normalise else 0, rows=_ptr(rows))
        if out.dtype == torch.bfloat16:
            g.out_bf16 = out.data_ptr()
        else:
            g.out_f32 = out.data_ptr()
        L.check(self.lib.pdlp_node_gather(C.byref(g), self.sh), "pdlp_node_gather (inputs)")

    def _gather_targets(self, Z: torch.Tensor, out: torch.Tensor, count: int, seed: int,
                        normalise: bool = True) -> None:
        sp = self.spec
        g = L.PdlpNodeGather(batch=Z.shape[0], nodes=sp.nodes, width=sp.channels, ld_out=self.d_out,
                             src_len=Z.shape[1], shared_len=0, count=count, seed=seed, src=Z.data_ptr(),
                             table=self.t_index.data_ptr(), scale=self.t_scale.data_ptr(),
                             mean=_ptr(self.out_mean) if normalise else 0,
                             inv_std=_ptr(self.out_inv_std) if normalise else 0, out_f32=out.data_ptr())
        L.check(self.lib.pdlp_node_gather(C.byref(g), self.sh), "pdlp_node_gather (targets)")

    def _upload(self, a) -> torch.Tensor:
        if isinstance(a, torch.Tensor):
            return a.contiguous()
        with torch.cuda.stream(self.mlp.stream):
            return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(self.device)

    # ---------------------------------------------------------------- forward
    def _forward_chunk(self, Vc: torch.Tensor) -> torch.Tensor:
        """Network output [rows, d_out] (FP32, the MLP's buffer for that row count) of
        every node of the instances Vc [b, src_len]."""
        P = None
        if self.prev is not None:
            P = self.prev._forward_chunk(Vc).view(Vc.shape[0], -1)
        rows = Vc.shape[0] * self.spec.nodes
        with torch.cuda.stream(self.mlp.stream):
            xb = torch.empty(rows, self.d_in, dtype=torch.bfloat16, device=self.device)
        self._gather_inputs(Vc, xb, P=P)
        return self.mlp.infer(xb)

    def node_outputs(self, Vd: torch.Tensor, chunk_rows: int = 1 << 20) -> torch.Tensor:
        """Normalised outputs of every node of every instance, [B, nodes * d_out] FP32
        on the device (the input of a second stage)."""
        B = Vd.shape[0]
        per = max(1, chunk_rows // self.spec.nodes)
        with torch.cuda.stream(self.mlp.stream):
            out = torch.empty(B, self.spec.nodes * self.d_out, dtype=torch.float32, device=self.device)
            for s in range(0, B, per):
                net = self._forward_chunk(Vd[s: s + per])
                out[s: s + per].copy_(net.view(-1, self.spec.nodes * self.d_out))
        return out

    # ---------------------------------------------------------------- training
    def _normalisers(self, Vd: torch.Tensor, Zd: torch.Tensor, sample: int, seed: int,
                     P: Optional[torch.Tensor] = None) -> None:
        st = self.mlp.stream
        with torch.cuda.stream(st):
            xs = torch.empty(sample, self.d_in, dtype=torch.float32, device=self.device)
            ts = torch.empty(sample, self.d_out, dtype=torch.float32, device=self.device)
        self._gather_inputs(Vd, xs, count=sample, seed=seed, normalise=False, P=P)
        self._gather_targets(Zd, ts, count=sample, seed=seed, normalise=False)
        st.synchronize()
        X, T = xs.cpu().numpy().astype(np.float64), ts.cpu().numpy().astype(np.float64)
        sp = self.spec
        w = self.width
        im = np.zeros(self.d_in)
        isd = np.ones(self.d_in)
        im[:w] = X[:, :w].mean(axis=0)
        isd[:w] = 1.0 / np.maximum(X[:, :w].std(axis=0), 1e-3)
        om = T[:, : sp.channels].mean(axis=0)
        osd = np.maximum(T[:, : sp.channels].std(axis=0), 1e-6)
        self.out_mean_host, self.out_std_host = om, osd
        with torch.cuda.stream(st):
            self.in_mean = torch.from_numpy(im.astype(np.float32)).to(self.device)
            self.in_inv_std = torch.from_numpy(isd.astype(np.float32)).to(self.device)
            self.out_mean = torch.from_numpy(np.pad(om, (0, self.d_out - sp.channels)).astype(np.float32)).to(self.device)
            self.out_inv_std = torch.from_numpy(np.pad(1.0 / osd, (0, self.d_out - sp.channels),
                                                       constant_values=1.0).astype(np.float32)).to(self.device)
        self._map = None

    def fit(self, V, Z, steps: int, rows: int = 16384, lr: float = 2e-3, warmup: float = 0.02, seed: int = 1,
            log=None, log_every: int = 500) -> list:
        """Train on per-instance vectors V [N, src_len] and solutions Z = [x*, y*]
        [N, n + m] (host arrays or device tensors) for ``steps`` Adam steps of
        ``rows`` (instance, node) rows drawn with replacement; the learning rate
        warms up linearly over the first ``warmup`` fraction, then follows a
        cosine to 0.  Returns the mean loss of each ``log_every`` steps."""
        t0 = time.perf_counter()
        Vd, Zd = self._upload(V), self._upload(Z)
        if Vd.shape[1] != self.spec.src_len or Zd.shape[1] != self.n + self.m or Vd.shape[0] != Zd.shape[0]:
            raise ValueError("V must be [N, src_len] and Z [N, n + m]")
        P = self.prev.node_outputs(Vd) if self.prev is not None else None
        self._normalisers(Vd, Zd, sample=min(1 << 18, Vd.shape[0] * self.spec.nodes), seed=0x5eed + seed, P=P)
        st = self.mlp.stream
        with torch.cuda.stream(st):
            xb = torch.empty(rows, self.d_in, dtype=torch.bfloat16, device=self.device)
            tb = torch.empty(rows, self.d_out, dtype=torch.float32, device=self.device)
        nw = max(1, int(warmup * steps))
        losses = []
        pending = []
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for s in range(steps):
            key = (seed << 32) + s + 1
            self._gather_inputs(Vd, xb, count=rows, seed=key, P=P)
            self._gather_targets(Zd, tb, count=rows, seed=key)
            loss = self.mlp.loss_and_grad(xb, tb, self.loss_weight)
            lr_s = lr * (s + 1) / nw if s < nw else lr * 0.5 * (1 + np.cos(np.pi * (s - nw) / max(1, steps - nw)))
            self.mlp.adam(lr=float(lr_s))
            if s % 50 == 0:
                with torch.cuda.stream(st):
                    pending.append(loss.clone())
            if (s + 1) % log_every == 0 or s + 1 == steps:
                st.synchronize()
                vals = [float(p.item()) for p in pending]
                pending = []
                losses.append(float(np.mean(vals)) if vals else float("nan"))
                if log:
                    log(f"step {s + 1}: loss {losses[-1]:.5f} ({time.perf_counter() - t0:.1f}s)")
        e1.record(st)
        st.synchronize()
        del P
        self.train_seconds = time.perf_counter() - t0
        # the step loop on the device (gathers, forward, backward, Adam), and its GEMM rate
        self.train_gpu_seconds = e0.elapsed_time(e1) * 1e-3
        d_h = self.mlp.d_h
        a, b, c = self.d_in * d_h, d_h * d_h, d_h * self.d_out
        # forward 2r(a+b+c); backward: dW of all three layers, dX of the upper two
        self.train_flop = 2.0 * rows * steps * (2 * a + 3 * b + 3 * c)
        return losses

    # ---------------------------------------------------------------- inference
    def _assembly(self):
        if self._map is None:
            if self.out_mean_host is None:
                raise RuntimeError("the node model is not trained")
            ln = self.n + self.m
            ip, nd, ch, w, b = assembly_map(self.spec, ln, self.out_mean_host, self.out_std_host)
            b = np.append(b, np.float32(0.0))                # entry n + m: log primal weight
            ip = np.append(ip, ip[-1]).astype(np.int32)
            with torch.cuda.stream(self.mlp.stream):
                up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(self.device)
                self._map = dict(indptr=up(ip), node=up(nd), chan=up(ch), weight=up(w), bias=up(b), nnz=len(nd))
        m = self._map
        lw = 0.0 if self.log_weight is None else float(self.log_weight)
        return m, lw

    def starts(self, V, chunk_rows: int = 1 << 20):
        """Start matrix [B, n + m + 1] (x, y, log weight; FP32, device) from per-instance
        vectors V [B, src_len] (host array or device tensor), and an event marking its
        completion on the network's stream."""
        Vd = self._upload(V)
        B = Vd.shape[0]
        m, lw = self._assembly()
        st = self.mlp.stream
        ln = self.n + self.m + 1
        with torch.cuda.stream(st):
            out = torch.empty(B, ln, dtype=torch.float32, device=self.device)
            m["bias"][-1:].fill_(lw)
        per = max(1, chunk_rows // self.spec.nodes)
        for s in range(0, B, per):
            Vc = Vd[s: s + per]
            bc = Vc.shape[0]
            net = self._forward_chunk(Vc)
            a = L.PdlpNodeAssembly(batch=bc, nodes=self.spec.nodes, ld_net=self.d_out, len=ln, ld_out=ln,
                                   nnz=m["nnz"], indptr=m["indptr"].data_ptr(), node=m["node"].data_ptr(),
                                   chan=m["chan"].data_ptr(), weight=m["weight"].data_ptr(),
                                   bias=m["bias"].data_ptr(), net=net.data_ptr(), out=out[s:].data_ptr())
            L.check(self.lib.pdlp_node_assemble(C.byref(a), self.sh), "pdlp_node_assemble")
        ev = torch.cuda.Event()
        ev.record(st)
        return out, ev

    def predict(self, V) -> tuple:
        """(x [B, n], y [B, m]) on the host, before projection (tests and analysis)."""
        out, ev = self.starts(V)
        ev.synchronize()
        o = out.cpu().numpy().astype(np.float64)
        return o[:, : self.n], o[:, self.n: self.n + self.m]

    def start_engine(self, eng, tol: float, V) -> None:
        """Start the family engine ``eng`` (loaded with the instances of V, in order) from
        the network's predictions."""
        out, ev = self.starts(V)
        n, m = self.n, self.m
        eng.start_from_device(tol, out, x_off=0, y_off=n, w_col=n + m if self.log_weight is not None else -1,
                              ready=ev)

    def calibrate_weight(self, fam, ids, tol: float, max_iters: int = 100_000,
                         candidates=(0.5, 1.0, 2.0, 4.0, 8.0), refine: bool = True, log=None) -> float:
        """Set the primal weight to the candidate (and the engine default, None) with the
        fewest mean iterations on instances ``ids`` (training instances held out of
        the fit's rows or not: the weight is one scalar), then (``refine``) try the
        best candidate times and divided by sqrt(2); returns it."""
        best, best_it = None, None

        def trial(w):
            nonlocal best, best_it
            self.log_weight = None if w is None else float(np.log(w))
            res, _ = solve_nodes(fam, ids, self, tol, max_iters)
            it = float(np.mean([r.iterations for r in res]))
            if log:
                log(f"weight {w if w is None else round(w, 3)}: mean iterations {it:.1f}")
            if best_it is None or it < best_it:
                best, best_it = self.log_weight, it

        for w in (None,) + tuple(candidates):
            trial(w)
        if refine and best is not None:
            w0 = float(np.exp(best))
            for w in (w0 / np.sqrt(2.0), w0 * np.sqrt(2.0)):
                trial(w)
        self.log_weight = best
        return best


def train_node_model(fam, lab, stages: int = 2, steps: int = 60_000, rows: int = 32768, lr: float = 2e-3,
                     d_h: int = 512, calib_ids=None, tol: float = 1e-4, max_iters: int = 100_000,
                     log=None) -> NodeModel:
    """Node model of ``stages`` stages (each reading the previous one's outputs at
    ``fam.node_neighbours()``) fitted to labels ``lab`` (:func:`.warmstart.label`;
    LPs that did not reach the tolerance are left out), each stage for ``steps``
    steps of ``rows`` rows; with ``calib_ids`` the primal weight is then calibrated
    on those instances.  Returns the last stage (which runs the ones before it)."""
    keep = np.asarray(lab.status) == "optimal"
    ids = np.asarray(lab.ids)[keep]
    V = fam.node_inputs(ids)
    Z = np.concatenate([lab.x[keep], lab.y[keep]], axis=1)
    spec = fam.node_spec()
    model = None
    for k in range(stages):
        model = NodeModel(spec, fam.n, fam.m, d_h=d_h, seed=k, prev=model,
                          neighbours=None if model is None else fam.node_neighbours())
        if k == 0:
            Vd, Zd = model._upload(V), model._upload(Z)
            del V, Z
        model.fit(Vd, Zd, steps=steps, rows=rows, lr=lr, seed=k + 1, log=log, log_every=max(1, steps // 4))
        if log:
            log(f"stage {k + 1}: {steps} steps in {model.train_seconds:.1f}s")
    del Vd, Zd
    if calib_ids is not None:
        model.calibrate_weight(fam, calib_ids, tol, max_iters, log=log)
    return model


def stage_list(model: NodeModel) -> list:
    """The stages of a node model, first to last."""
    out = []
    while model is not None:
        out.append(model)
        model = model.prev
    return out[::-1]


def forward_flop(model: NodeModel, rows: int) -> float:
    """FLOP of one forward pass of every stage over ``rows`` node rows (the three GEMMs)."""
    return sum(2.0 * rows * (s.d_in * s.mlp.d_h + s.mlp.d_h * s.mlp.d_h + s.mlp.d_h * s.d_out)
               for s in stage_list(model))


def solve_nodes(fam, ids, model: NodeModel, tol: float, max_iters: int, chunk: int = 2048,
                scheme: str = "adaptive", engines: Optional[dict] = None):
    """Node-model-warm-started batched solves of family instances; returns SolveResults
    and the seconds from the per-instance vectors on the host to the results on the
    host (upload, network and assembly included)."""
    res = []
    secs = 0.0
    engines = {} if engines is None else engines
    for s in range(0, len(ids), chunk):
        part = ids[s: s + chunk]
        eng = _engine_for(engines, [fam.problem(int(k)) for k in part], tol, scheme)
        V = fam.node_inputs(part)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        model.start_engine(eng, tol, V)
        eng.run(max_iters)
        o, xs, ys = eng.results()
        secs += time.perf_counter() - t0
        res += results_of(eng, o, xs, ys)
    return res, secs
