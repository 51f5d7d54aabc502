"""Host logic of the autotuner: the coordinate search, the cache, the shape buckets
and the persistent packer's FP32/FP64 split (no device)."""

import ctypes as C
import json

import numpy as np

from sovereign_solver_amd.engine import _lib as L
from sovereign_solver_amd.engine import tune as T
from sovereign_solver_amd.engine.batch import PackedBatch, lds_bytes_resident
from sovereign_solver_amd.workloads import uneven_batch


def test_coordinate_search_keeps_the_fastest_optimal_trial():
    # seconds = a separable bowl with its minimum at (b=2, c=30); c=40 is fastest but fails
    def measure(cfg):
        ok = cfg["c"] != 40
        s = (cfg["b"] - 2) ** 2 + abs(cfg["c"] - 30) / 10 + (0 if ok else -100)
        return {"seconds": s, "ok": ok}
    best, trials = T.coordinate_search({"b": (1, 2, 3), "c": (10, 30, 40)}, {"b": 1, "c": 10}, measure)
    assert best == {"b": 2, "c": 30}
    assert all("ok" in v for v in trials.values())
    # one pass: each value of each parameter once (the shared start is measured once)
    assert len(trials) == 5


def test_coordinate_search_without_an_optimal_trial_keeps_the_start():
    best, _ = T.coordinate_search({"a": (1, 2)}, {"a": 7}, lambda cfg: {"seconds": 1.0, "ok": False})
    assert best == {"a": 7}


def test_cache_round_trip(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    monkeypatch.setenv("SOVEREIGN_PDLP_TUNE_CACHE", str(path))
    t = T.Tune(4, 256, 256, 100, 1, 0, 0, 1, 6, 0)
    T._store_cache(T._bucket(3000, 2_000_000, 4_000_000, "fp32"), t, {"primal/tpr4": 15.0})
    assert T.cached_tune(3000, 2_000_000, 4_000_000, "fp32", check_every=None) == t
    assert T.cached_tune(3000, 2_000_000, 4_000_000, "fp32", check_every=50).check_every == 50
    assert T.cached_tune(3000, 2_000_000, 8_000_000, "fp32") is None          # another bucket
    T._store_entry("gfx950/persistent/B12", {"config": {"block": 512}})
    data = json.loads(path.read_text())
    assert set(data) == {T._bucket(3000, 2_000_000, 4_000_000, "fp32"), "gfx950/persistent/B12"}


def test_batch_bucket_separates_shapes():
    a = T.batch_bucket("persistent", 4096, [50, 100, 2000], [500, 1000, 20000])
    b = T.batch_bucket("persistent", 4096, [50, 100, 4000], [500, 1000, 20000])
    c = T.batch_bucket("persistent", 4096, [50, 100, 2000], [500, 1000, 20000], "scaled")
    assert len({a, b, c}) == 3 and a.startswith("gfx950/persistent/B12/")


def test_split_moves_lps_between_the_parts():
    pk = PackedBatch(uneven_batch(64, seed=5))
    need = lds_bytes_resident(pk.ms, pk.ns)
    for split in (0, 8 * 1024, 16 * 1024, 32 * 1024, 1 << 30):
        pk.set_split(split)
        res = need <= split
        assert pk.n_resident == int(res.sum())
        # the FP32 part first, then the FP64 part, each largest (by entries) first
        first, second = pk.order[: len(pk.order) - pk.n_resident], pk.order[len(pk.order) - pk.n_resident:]
        assert not res[first].any() and res[second].all()
        for part in (first, second):
            assert np.all(np.diff(pk.nnzs[part]) <= 0)
        assert sorted(pk.order.tolist()) == list(range(64))
        if pk.n_resident:
            assert pk.lds_need_resident == int(need[res].max()) <= pk.lds_resident
        if pk.n_resident < 64:
            assert 0 < pk.lds_need <= pk.lds_bytes


def test_batch_rejects_lines_outside_two_to_four_without_device():
    from tests.test_abi import _batch_args
    lib = L.lib()
    a, _buf = _batch_args()
    a.lds_need = 512
    a.lines = 5
    assert lib.pdlp_batch_solve(C.byref(a), None) == L.PDLP_E_TUNE
    assert b"lines" in lib.pdlp_last_error()
