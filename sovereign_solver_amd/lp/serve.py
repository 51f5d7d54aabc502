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

from ..config import MAX_PDHG_ITERS
from ..engine.family import FamilyBatch, FamilyEngine
from ..problem import Problem


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
