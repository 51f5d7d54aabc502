"""Launch-shape and check-cadence selection for the PDHG engines.

Layers:

* :func:`heuristic_tune` — lanes per row from the mean row length of K (dual
  step) and Kᵀ (primal step): the largest power of two not above the mean,
  within 1..64.  Short rows (config-1's Kᵀ has 2 entries per row) get one or
  two lanes; long rows (its K rows have 1,000–2,000) get a full wavefront.
* :func:`autotune` — times each candidate of a small grid around the
  heuristic on the actual problem with HIP events on the engine's stream and
  keeps the fastest; results are cached per (arch, shape bucket) in a JSON
  file (``SOVEREIGN_PDLP_TUNE_CACHE`` or ``~/.cache/sovereign_pdlp/tune.json``).
* :func:`tune_cadence` — the restart-check cadence of the single-instance
  engine by measured time to tolerance: the LP is solved once per candidate
  cadence (a check every 25, 50, 100 or 200 iterations) and the fastest
  solve that ends optimal wins.  The cadence trades the check's cost (two
  FP64 passes over K per candidate point) against how late a solve notices
  convergence and how often the restart rule looks, so only a whole solve
  measures it.
* :func:`tune_family` and :func:`tune_batch` — the same for the
  shared-matrix family engine (cadence) and the persistent batch engine
  (cadence, workgroup size of each part, lines per lane group and round of
  the FP32 part, and the FP32/FP64 split), by a coordinate search over
  whole batched solves.

Every choice is cached per shape bucket (:func:`_bucket`, :func:`batch_bucket`),
with the timings that decided it.
"""

from __future__ import annotations

import json
import math
import os
import threading
import time
from dataclasses import asdict, dataclass
from typing import Callable, Optional

ARCH = "gfx950"
# Version of the single-instance launch choices stored in the cache (raised when a
# stored field changes meaning: 2 = tpr_dual sets the CSR rows' threads in SELL modes 2..7)
SINGLE_LAYOUT = 2
TPR_CHOICES = (1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024)
CADENCES = (25, 50, 100, 200)          # iterations between restart checks


@dataclass(frozen=True)
class Tune:
    tpr_primal: int = 1
    tpr_dual: int = 1
    block: int = 256
    check_every: int = 50
    xcd_remap: int = 0
    exact: int = 0
    block_dual: int = 0
    scheme: int = 0             # 0 reference restart rule, 1 adaptive (see include/pdlp_hip.h)
    sell: int = 0               # SELL-64 dual step: 0 off, 1 a wavefront per slice, 2..7 one launch with a
                                # workgroup per slice piece (threads/rows: 2 1024/64, 3 512/64, 4 1024/16,
                                # 5 1024/8, 6 1024/32, 7 512/16) and tpr_dual (>= 64) threads per CSR row,
                                # k >= 8 a wavefront per k entries
    no_ray: int = 0             # 1: the adaptive checks skip the unboundedness test
    sell_primal: int = 0        # primal step over the SELL-64 copy of K^T: 1 one slice, 2 two slices per wavefront

    def to_c(self):
        from ._lib import PdlpTune
        return PdlpTune(self.tpr_primal, self.tpr_dual, self.block, self.check_every, self.xcd_remap, self.exact,
                        self.block_dual, self.scheme, self.sell, self.no_ray, self.sell_primal)

    def with_(self, **kw) -> "Tune":
        d = asdict(self)
        d.update(kw)
        return Tune(**d)


def _tpr_for(mean_len: float) -> int:
    """Lanes per row: about one lane per 2–4 entries, a workgroup per row above 256."""
    if mean_len <= 1.0:
        return 1
    t = 1 << int(math.floor(math.log2(mean_len)))
    if t > 64:                     # long rows: a workgroup per row, ~4 entries per lane
        return max(128, min(1024, t // 4))
    return max(1, t)


def heuristic_tune(m: int, n: int, nnz: int, check_every: int = 50, exact: bool = False) -> Tune:
    if exact:
        return Tune(1, 1, 256, check_every, 0, 1)
    return Tune(tpr_primal=_tpr_for(nnz / max(n, 1)), tpr_dual=_tpr_for(nnz / max(m, 1)),
                block=256, check_every=check_every, xcd_remap=0, exact=0)


def _bucket(m: int, n: int, nnz: int, prec: str) -> str:
    lg = lambda v: int(round(math.log2(max(v, 1))))
    return f"{ARCH}/{prec}/s{SINGLE_LAYOUT}/m{lg(m)}/n{lg(n)}/z{lg(nnz)}"


def _cache_path() -> str:
    return os.environ.get("SOVEREIGN_PDLP_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "sovereign_pdlp", "tune.json"))


_cache_lock = threading.Lock()


def _load_cache() -> dict:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(key: str, t: Tune, timings: dict) -> None:
    _store_entry(key, {"tune": asdict(t), "timings_us": timings})


def _store_entry(key: str, entry: dict) -> None:
    with _cache_lock:
        data = _load_cache()
        data[key] = entry
        path = _cache_path()
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            tmp = path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(data, f, indent=1, sort_keys=True)
            os.replace(tmp, path)
        except OSError:
            pass


def cached_tune(m: int, n: int, nnz: int, prec: str, check_every: Optional[int] = 50):
    """The cached launch shape of the bucket; ``check_every=None`` keeps the cached
    cadence (the one :func:`tune_cadence` chose, when it ran), else the given one."""
    rec = _load_cache().get(_bucket(m, n, nnz, prec))
    if not rec:
        return None
    d = dict(rec["tune"])
    if check_every is not None:
        d["check_every"] = check_every
    return Tune(**d)


def _neighbours(t: int):
    i = TPR_CHOICES.index(t)
    return sorted({TPR_CHOICES[k] for k in range(max(0, i - 2), min(len(TPR_CHOICES), i + 3))})


def autotune(engine, reps: int = 20, use_cache: bool = True) -> Tune:
    """Pick tpr_primal / tpr_dual / xcd_remap for ``engine`` by timing its kernels.

    The engine's state is overwritten (it is re-initialised by the next solve).
    """
    key = _bucket(engine.m, engine.n, engine.nnz, engine.precision)
    if use_cache:
        hit = cached_tune(engine.m, engine.n, engine.nnz, engine.precision, engine.tune.check_every)
        if hit is not None:
            return hit.with_(scheme=engine.tune.scheme)
    base = heuristic_tune(engine.m, engine.n, engine.nnz, engine.tune.check_every)
    timings = {}
    best_p, best_pt = None, float("inf")
    for tp in _neighbours(base.tpr_primal):
        for remap in (0, 1):
            t = Tune(tp, base.tpr_dual, 256, base.check_every, remap, 0)
            us = engine.time_kernel("primal", t, reps)
            timings[f"primal/tpr{tp}/remap{remap}"] = us
            if us < best_pt:
                best_p, best_pt = (tp, remap), us
    # the SELL-64 copy of K^T (one lane per column, no column pointers), when the engine has one
    sell_primal, block_p = 0, 256
    if getattr(engine, "t_sell_slices", 0) > 0:
        for sp in (1, 2):                 # one or two slices per wavefront
            for bp in (256, 512, 1024):
                for remap in (0, 1):
                    t = Tune(1, base.tpr_dual, bp, base.check_every, remap, 0, sell_primal=sp)
                    us = engine.time_kernel("primal", t, reps)
                    timings[f"primal/sell{sp}/block{bp}/remap{remap}"] = us
                    if us < best_pt:
                        best_p, best_pt, sell_primal, block_p = (1, remap), us, sp, bp
    best_d, best_dt = None, float("inf")
    for td in _neighbours(base.tpr_dual):
        # several rows per workgroup share gathered cache lines in L1 when rows are short enough
        blocks = (256, 1024) if td <= 64 else (0,)
        for bd in blocks:
            for remap in (0, 1):
                t = Tune(best_p[0], td, 256, base.check_every, remap << 1, 0, bd)
                us = engine.time_kernel("dual", t, reps)
                timings[f"dual/tpr{td}/block{bd}/remap{remap}"] = us
                if us < best_dt:
                    best_d, best_dt = (td, remap, bd), us
    sell = 0
    if getattr(engine, "sell_slices", 0) > 0:
        # the SELL slices take the coalescing rows; the CSR kernel keeps the rest
        for sv in (1, 2, 3, 4, 5, 6, 7, 16, 32, 64, 128):
            # one-launch modes: threads per CSR row (no remap); others: the CSR kernel's lanes
            one = 2 <= sv <= 7
            tds = [w for w in (64, 128, 256, 512, 1024) if w <= (512 if sv in (3, 7) else 1024)] if one \
                else _neighbours(base.tpr_dual)
            for td in tds:
                for remap in ((0,) if one else (0, 1)):
                    t = Tune(best_p[0], td, 256, base.check_every, remap << 1, 0, 0, 0, sv)
                    us = engine.time_kernel("dual", t, reps)
                    timings[f"dual/sell{sv}/tpr{td}/remap{remap}"] = us
                    if us < best_dt:
                        best_d, best_dt, sell = (td, remap, 0), us, sv
    # `block` is the primal step's (the dual step has block_dual, 0 = 256 here)
    t = Tune(best_p[0], best_d[0], block_p, base.check_every, best_p[1] | (best_d[1] << 1), 0,
             best_d[2] or (256 if block_p != 256 and best_d[0] <= 64 else 0), 0, sell, sell_primal=sell_primal)
    _store_cache(key, t, timings)
    return t.with_(scheme=engine.tune.scheme)


# ------------------------------------------------------------------ cadence
def _fastest(trials: dict) -> Optional[object]:
    """Key of the fastest trial that ended with every LP optimal (None if none did)."""
    ok = {k: v for k, v in trials.items() if v["ok"]}
    return min(ok, key=lambda k: ok[k]["seconds"]) if ok else None


def tune_cadence(engine, tol: float, max_iters: int, candidates=CADENCES, use_cache: bool = True,
                 x0=None, y0=None, step_scale=None) -> Tune:
    """Check cadence of a :class:`~.single.SingleEngine` by time to tolerance.

    Solves the engine's LP once per candidate cadence (from the same start) and
    keeps the fastest solve that ends ``optimal``; ties and failures keep the
    current cadence.  The choice and the timings are stored with the bucket's
    launch shape.  Returns the tuning with the chosen cadence (also set on the
    engine)."""
    key = _bucket(engine.m, engine.n, engine.nnz, engine.precision)
    rec = _load_cache().get(key) if use_cache else None
    if rec and "cadence" in rec:
        t = engine.tune.with_(check_every=int(rec["cadence"]["check_every"]))
        engine.set_tune(t)
        return t
    base = engine.tune
    trials = {}
    for ce in candidates:
        engine.set_tune(base.with_(check_every=int(ce)))
        engine.stream.synchronize()
        t0 = time.perf_counter()
        r = engine.solve(tol, max_iters, x0, y0, step_scale=step_scale)
        engine.stream.synchronize()
        trials[int(ce)] = {"seconds": time.perf_counter() - t0, "ok": r["status"] == "optimal",
                           "iterations": int(r["iterations"])}
    best = _fastest(trials)
    ce = best if best is not None else base.check_every
    t = base.with_(check_every=int(ce))
    engine.set_tune(t)
    entry = dict(rec or _load_cache().get(key) or {"tune": asdict(base), "timings_us": {}})
    entry["cadence"] = {"check_every": int(ce), "tol": tol,
                        "trials": {str(k): v for k, v in trials.items()}}
    _store_entry(key, entry)
    return t


def coordinate_search(space: dict, start: dict, measure: Callable[[dict], dict]):
    """One pass of coordinate descent: for each parameter in ``space`` order, try
    each of its values with the others at the best so far, keep the fastest
    trial whose LPs all ended optimal.  ``measure(config)`` returns
    ``{"seconds": s, "ok": bool, ...}``.  Returns (best config, trials)."""
    best = dict(start)
    trials = {}

    def key(cfg):
        return ",".join(f"{k}={cfg[k]}" for k in sorted(cfg))

    best_rec = None
    for name, values in space.items():
        for v in values:
            cfg = dict(best)
            cfg[name] = v
            kk = key(cfg)
            if kk not in trials:
                trials[kk] = measure(cfg)
            rec = trials[kk]
            if rec["ok"] and (best_rec is None or rec["seconds"] < best_rec["seconds"]):
                best, best_rec = cfg, rec
    return best, trials


def _lg(v) -> int:
    return int(round(math.log2(max(float(v), 1.0))))


def batch_bucket(kind: str, count: int, ms, nnzs, extra: str = "") -> str:
    """Shape bucket of a batch: engine kind, log2 of the LP count, of the median and
    largest row count and of the total entries."""
    import numpy as np
    ms = np.asarray(ms)
    return (f"{ARCH}/{kind}/B{_lg(count)}/m{_lg(np.median(ms))}/M{_lg(ms.max())}/z{_lg(np.sum(nnzs))}"
            + (f"/{extra}" if extra else ""))


def tune_family(engine, tol: float, max_iters: int, X=None, Y=None, w=None, candidates=(25, 50, 100),
                use_cache: bool = True) -> dict:
    """Check cadence of a :class:`~.family.FamilyEngine` by time to tolerance of the
    whole batch (from the given starts); sets it on the engine and returns the
    cached record ``{"check_every": ..., "trials": ...}``."""
    fb = engine.fb
    key = batch_bucket("family", fb.count, [fb.m], [fb.K.nnz], f"n{_lg(fb.n)}")
    rec = _load_cache().get(key) if use_cache else None
    if rec is None:
        def measure(cfg):
            engine.set_check_every(cfg["check_every"])
            engine.stream.synchronize()
            t0 = time.perf_counter()
            out, _, _ = engine.solve(tol, max_iters, X, Y, w)
            return {"seconds": time.perf_counter() - t0, "ok": bool((out["status"] == 1).all()),
                    "mean_iterations": float(out["iterations"].mean())}
        best, trials = coordinate_search({"check_every": tuple(candidates)},
                                         {"check_every": engine.check_every}, measure)
        rec = {"check_every": int(best["check_every"]), "tol": tol, "trials": trials}
        _store_entry(key, rec)
    engine.set_check_every(int(rec["check_every"]))
    return rec


BATCH_SPACE = {
    "check_every": (25, 50, 100),
    "block": (256, 512),                # FP32 part: threads per workgroup
    "lines": (2, 3, 4),                 # FP32 part: lines per lane group and round
    "split": (8 * 1024, 16 * 1024, 32 * 1024),   # FP64 part: LPs needing at most this much LDS
    "block_resident": (256, 512),       # FP64 part: threads per workgroup
}
BATCH_DEFAULT = {"check_every": 50, "block": 512, "lines": 3, "split": 16 * 1024, "block_resident": 256}


def tune_batch(engine, tol: float, max_iters: int, x0=None, y0=None, w=None, space=None,
               use_cache: bool = True) -> dict:
    """Launch configuration of a :class:`~.batch.BatchEngine` by time to tolerance of
    the whole batch: a coordinate search over ``space`` (default
    :data:`BATCH_SPACE`) from :data:`BATCH_DEFAULT`, each trial one batched solve
    from the given start, timed with HIP events on the engine stream.  The best
    configuration is set on the engine; returns ``{"config": ..., "trials": ...}``."""
    pk = engine.pk
    key = batch_bucket("persistent", len(pk.problems), pk.ms, pk.nnzs,
                       "scaled" if pk.row_scale is not None else "")
    rec = _load_cache().get(key) if use_cache else None
    if rec is None:
        import torch

        engine.set_warm_start(x0, y0, w)              # kernels loaded before the first timed trial
        engine.launch(tol, min(max_iters, 1), engine.check_every)
        engine.stream.synchronize()

        def measure(cfg):
            engine.configure(**cfg)
            engine.set_warm_start(x0, y0, w)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(engine.stream)
            engine.launch(tol, max_iters, engine.check_every)
            e1.record(engine.stream)
            out, _, _ = engine.results()
            return {"seconds": e0.elapsed_time(e1) * 1e-3, "ok": bool((out["status"] == 1).all()),
                    "max_iterations": int(out["iterations"].max())}
        best, trials = coordinate_search(space or BATCH_SPACE, BATCH_DEFAULT, measure)
        rec = {"config": best, "tol": tol, "trials": trials}
        _store_entry(key, rec)
    engine.configure(**rec["config"])
    return rec
