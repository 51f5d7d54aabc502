"""Launch-shape selection for the PDHG kernels.

Two layers:

* :func:`heuristic_tune` — lanes per row from the mean row length of K (dual
  step) and Kᵀ (primal step): the largest power of two not above the mean,
  within 1..64.  Short rows (config-1's Kᵀ has 2 entries per row) get one or
  two lanes; long rows (its K rows have 1,000–2,000) get a full wavefront.
* :func:`autotune` — times each candidate of a small grid around the
  heuristic on the actual problem with HIP events on the engine's stream and
  keeps the fastest; results are cached per (arch, shape bucket) in a JSON
  file (``SOVEREIGN_PDLP_TUNE_CACHE`` or ``~/.cache/sovereign_pdlp/tune.json``).
"""

from __future__ import annotations

import json
import math
import os
import threading
from dataclasses import asdict, dataclass

ARCH = "gfx950"
TPR_CHOICES = (1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024)


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
                                # workgroup per CSR row and per slice piece (threads/rows: 2 1024/64, 3 512/64,
                                # 4 1024/16, 5 1024/8, 6 1024/32, 7 512/16), k >= 8 a wavefront per k entries
    no_ray: int = 0             # 1: the adaptive checks skip the unboundedness test

    def to_c(self):
        from ._lib import PdlpTune
        return PdlpTune(self.tpr_primal, self.tpr_dual, self.block, self.check_every, self.xcd_remap, self.exact,
                        self.block_dual, self.scheme, self.sell, self.no_ray)

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
    return f"{ARCH}/{prec}/m{lg(m)}/n{lg(n)}/z{lg(nnz)}"


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
    with _cache_lock:
        data = _load_cache()
        data[key] = {"tune": asdict(t), "timings_us": timings}
        path = _cache_path()
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            tmp = path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(data, f, indent=1, sort_keys=True)
            os.replace(tmp, path)
        except OSError:
            pass


def cached_tune(m: int, n: int, nnz: int, prec: str, check_every: int = 50):
    rec = _load_cache().get(_bucket(m, n, nnz, prec))
    if not rec:
        return None
    d = dict(rec["tune"])
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
            for td in _neighbours(base.tpr_dual):
                for remap in (0, 1):
                    t = Tune(best_p[0], td, 256, base.check_every, remap << 1, 0, 0, 0, sv)
                    us = engine.time_kernel("dual", t, reps)
                    timings[f"dual/sell{sv}/tpr{td}/remap{remap}"] = us
                    if us < best_dt:
                        best_d, best_dt, sell = (td, remap, 0), us, sv
    t = Tune(best_p[0], best_d[0], 256, base.check_every, best_p[1] | (best_d[1] << 1), 0, best_d[2], 0, sell)
    _store_cache(key, t, timings)
    return t.with_(scheme=engine.tune.scheme)
