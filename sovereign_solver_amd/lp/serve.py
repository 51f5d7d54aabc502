"""Serving loop over consecutive batches of one LP family, host to host.

A caller that solves a stream of batches (branch-and-bound rounds, scenario
sweeps, the held-out batches of a learned warm start) spends, per batch, host
time that the device does not need: building the network's inputs, gathering
the per-LP vectors into pinned staging, reading results back.  One engine
would leave the GPU idle meanwhile.  :class:`FamilyServer` keeps two family
engines of the same shape and prepares batch k+1 on a worker thread (inputs,
staging, upload, ``pdlp_family_load``, network inference on the network's
stream) while batch k runs on the device; the main thread then starts batch
k+1 from the prepared state as soon as batch k's results are out.

The reference has no batch call (its callers loop over
``solver/lp/pdhg.py:152``); each batch's results have the semantics of
:meth:`~sovereign_solver_amd.engine.family.FamilyEngine.results`.
"""

from __future__ import annotations

import concurrent.futures as cf
from typing import Callable, Iterable, Iterator, Optional, Sequence

import torch

import numpy as np

from ..config import MAX_PDHG_ITERS
from ..engine.batch import OUT_DTYPE
from ..engine.family import FamilyBatch, FamilyEngine, StreamBatch
from ..problem import Problem


class _Staged:
    """What a features function reads of an engine (:meth:`FamilyEngine.staged` and
    its stream), for sources uploaded on their own stream."""

    def __init__(self, src, stream):
        self._src, self.stream = src, stream

    def staged(self, name):
        return self._src.get(name)


class FamilyServer:
    """Two :class:`FamilyEngine` of one family and batch size, used alternately.

    ``problems``: a batch of the family (its matrices, sharing pattern and size
    fix the engines; every later batch must match, see
    :meth:`FamilyBatch.with_problems`).  ``model`` (optional): a network with
    ``starts(V) -> (start matrix, event)``, its stream ``mlp.stream`` and ``n``,
    ``m``, ``log_weight`` (:class:`~sovereign_solver_amd.learned.nodes.NodeModel`),
    fed by ``features(problems, engine) -> V``, called after the batch is loaded
    into ``engine``: a host array, or a device tensor written on the engine's
    stream (e.g. :meth:`~sovereign_solver_amd.workloads.DispatchFamily.node_inputs_staged`);
    without a model every batch starts cold.
    ``engine_kw``: launch shape and cadence of the engines (``items_per_wave``,
    ``xcd_affinity``, ``check_every``, ``scheme``, ...)."""

    def __init__(self, problems: Sequence[Problem], tol: float = 1e-4, max_iters: int = MAX_PDHG_ITERS,
                 model=None, features: Optional[Callable] = None, engine_kw: Optional[dict] = None):
        if (model is None) != (features is None):
            raise ValueError("a model needs its features function, and features a model")
        fb = FamilyBatch(problems)
        kw = dict(engine_kw or {})
        self.engines = [FamilyEngine(fb, **kw), FamilyEngine(fb.with_problems(problems), **kw)]
        self.tol, self.max_iters = float(tol), int(max_iters)
        self.model, self.features = model, features
        self.count = fb.count
        self._up = torch.cuda.Stream(device=self.engines[0].device)   # uploads of streamed batches

    def _prepare(self, k: int, problems: Sequence[Problem]):
        """Host and upload work of batch k (worker thread): staging and
        ``pdlp_family_load`` on engine k % 2, then the network's starts."""
        eng = self.engines[k % 2]
        eng.load(problems)
        if self.model is None:
            return eng, None, None
        V = self.features(problems, eng)
        if isinstance(V, torch.Tensor):
            ready = torch.cuda.Event()
            ready.record(eng.stream)
            self.model.mlp.stream.wait_event(ready)
            V.record_stream(self.model.mlp.stream)     # read there after this frame drops it
        out, ev = self.model.starts(V)
        return eng, out, ev

    def _finish(self, prep):
        """Device solve of a prepared batch and its results (main thread)."""
        eng, out, ev = prep
        if out is None:
            eng.start(self.tol)
        else:
            n, m = self.model.n, self.model.m
            w_col = n + m if self.model.log_weight is not None else -1
            eng.start_from_device(self.tol, out, x_off=0, y_off=n, w_col=w_col, ready=ev)
        eng.run(self.max_iters)
        o, X, Y = eng.results()
        return o, X, Y, eng.objectives(X)

    def _stream_batch(self, problems: Sequence[Problem]):
        """A batch as a :class:`StreamBatch` (worker thread): its sources gathered into
        fresh pinned memory and uploaded on the server's upload stream, then, with a
        model, its network inputs and starts; returns (batch, host copy of c or None)."""
        eng = self.engines[0]
        eng.fb.with_problems(problems)                 # same family, count and sharing
        host = eng.host_sources(problems)
        with torch.cuda.stream(self._up):
            src = {k: h.to(eng.device, non_blocking=True) for k, h in host.items()}
            up = torch.cuda.Event()
            up.record(self._up)
        if self.model is None:
            return StreamBatch(src, ready=up, count=len(problems)), host.get("c")
        V = self.features(problems, _Staged(src, self._up))
        ready = torch.cuda.Event()                     # the upload and device-side inputs
        ready.record(self._up)
        self.model.mlp.stream.wait_event(ready)        # so the starts' event covers them too
        if isinstance(V, torch.Tensor):
            V.record_stream(self.model.mlp.stream)
        out, ev = self.model.starts(V)
        n, m = self.model.n, self.model.m
        w_col = n + m if self.model.log_weight is not None else -1
        return StreamBatch(src, start=out, ready=ev, x_off=0, y_off=n, w_col=w_col), host.get("c")

    def solve_streamed(self, batches: Sequence[Sequence[Problem]]):
        """Every batch through one engine as one stream (:meth:`FamilyEngine.run_stream`):
        a worker thread gathers, uploads and starts batch k+1 while the engine runs,
        and the LPs of a batch enter tiles as earlier LPs finish, so no batch waits on
        another's slowest LPs.  Returns [(out records, X, Y, objectives)] per batch,
        host arrays of the caller's own; each LP's result is bitwise :meth:`solve`'s
        on its batch.  ``max_iters`` must be a multiple of the engines' check_every."""
        batches = [list(b) for b in batches]
        if not batches:
            return []
        eng = self.engines[0]
        if self.max_iters % eng.check_every:
            raise ValueError("a stream's max_iters must be a multiple of check_every")
        total = sum(len(b) for b in batches)
        costs = []
        with cf.ThreadPoolExecutor(1) as ex:
            futs = [ex.submit(self._stream_batch, probs) for probs in batches]

            def source():
                for f in futs:
                    sb, c_host = f.result()
                    costs.append(c_host)
                    yield sb
            try:
                O, X, Y = eng.run_stream(self.tol, self.max_iters, source(), total=total)
            except BaseException:
                for f in futs:
                    f.cancel()
                cf.wait(futs)
                raise
        with torch.cuda.stream(eng.stream):
            o_all = np.frombuffer(O.cpu().numpy().tobytes(), dtype=OUT_DTYPE)
            Xh, Yh = X.cpu().numpy(), Y.cpu().numpy()
        res, lo = [], 0
        for probs, c_host in zip(batches, costs):
            hi = lo + len(probs)
            Xb = Xh[lo:hi]
            obj = (np.einsum("ij,ij->i", c_host.numpy(), Xb) if c_host is not None
                   else Xb @ np.asarray(probs[0].c, float))
            obj = obj * np.array([-1.0 if p.sense == "max" else 1.0 for p in probs])
            res.append((o_all[lo:hi].copy(), Xb.copy(), Yh[lo:hi].copy(), obj))
            lo = hi
        return res

    def solve(self, problems: Sequence[Problem]):
        """One batch, without overlap: (out records, X [B, n], Y [B, m], objectives)."""
        return self._finish(self._prepare(0, problems))

    def solve_stream(self, batches: Iterable[Sequence[Problem]]) -> Iterator[tuple]:
        """(out records, X, Y, objectives) per batch, in order.  X and Y are views of
        an engine's pinned buffers: use (or copy) them before asking for the next
        batch's results, which may reload that engine."""
        it = iter(batches)
        first = next(it, None)
        if first is None:
            return
        with cf.ThreadPoolExecutor(1) as ex:
            fut = ex.submit(self._prepare, 0, first)
            k = 0
            while True:
                prep = fut.result()
                nxt = next(it, None)
                if nxt is not None:
                    # batch k+1 goes to the other engine, idle since its results were read
                    fut = ex.submit(self._prepare, k + 1, nxt)
                try:
                    res = self._finish(prep)
                except BaseException:
                    if nxt is not None:
                        fut.cancel()
                        cf.wait([fut])
                    raise
                yield res
                if nxt is None:
                    return
                k += 1
