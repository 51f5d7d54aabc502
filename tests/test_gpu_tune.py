"""GPU: the autotuner's choices still solve to the tolerance.

Each engine is tuned on its own batch (a private cache file), and the LPs it then
solves with the chosen configuration are judged on the host: the reference KKT
formula (oracle yardstick, FP64) at the requested tolerance for every LP.  The
persistent engine's lines per round change no arithmetic (each line's sum runs
over the same lanes in the same order), so they give the same bits as the
default; its workgroup sizes change the order of the check's block sums.
"""

import numpy as np
import pytest

from oracle import yardstick
from sovereign_solver_amd import solve_lp_pdhg, solve_lp_pdhg_batch
from sovereign_solver_amd.engine import tune as T
from sovereign_solver_amd.engine.batch import BatchEngine, PackedBatch
from sovereign_solver_amd.workloads import DispatchFamily, random_sparse_lp, transportation_lp, uneven_batch

pytestmark = pytest.mark.gpu
TOL = 1e-4


@pytest.fixture
def cache(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    monkeypatch.setenv("SOVEREIGN_PDLP_TUNE_CACHE", str(path))
    return path


def _all_within(probs, res, tol=TOL):
    for p, r in zip(probs, res):
        assert r.status == "optimal", (p.name, r.status)
        e = yardstick.rel_kkt(p, r.x, r.y)[0]
        assert e <= 1.02 * tol, (p.name, e)


def test_single_instance_cadence_tuned_solve(gpu, cache):
    p = transportation_lp(60, 200, seed=4)
    r = solve_lp_pdhg(p, tol=TOL, precision="fp32", tune="auto")
    _all_within([p], [r])
    rec = T._load_cache()[next(iter(T._load_cache()))]
    assert rec["cadence"]["check_every"] in T.CADENCES
    trials = rec["cadence"]["trials"]
    assert set(map(int, trials)) == set(T.CADENCES)
    # the chosen cadence is the fastest optimal trial
    ok = {int(k): v["seconds"] for k, v in trials.items() if v["ok"]}
    assert rec["cadence"]["check_every"] == min(ok, key=ok.get)
    # the second call takes the cached choice and solves the same way
    r2 = solve_lp_pdhg(p, tol=TOL, precision="fp32", tune="auto")
    _all_within([p], [r2])
    assert r2.iterations == r.iterations


def test_persistent_batch_tuned_solve(gpu, cache):
    probs = uneven_batch(96, seed=7)
    res = solve_lp_pdhg_batch(probs, tol=TOL, engine="persistent", tune=True)
    _all_within(probs, res)
    rec = next(v for k, v in T._load_cache().items() if "/persistent/" in k)
    cfg = rec["config"]
    assert cfg["block"] in (256, 512) and cfg["lines"] in (2, 3, 4) and cfg["block_resident"] in (256, 512)
    assert cfg["check_every"] in T.BATCH_SPACE["check_every"] and cfg["split"] in T.BATCH_SPACE["split"]
    chosen = ",".join(f"{k}={cfg[k]}" for k in sorted(cfg))
    assert rec["trials"][chosen]["ok"]


@pytest.mark.parametrize("lines", [2, 4])
def test_persistent_lines_per_round_change_no_bits(gpu, lines):
    probs = uneven_batch(48, seed=9)
    pk = PackedBatch(probs)
    base = BatchEngine(pk).solve(TOL, 50_000)
    eng = BatchEngine(pk)
    eng.configure(lines=lines)
    other = eng.solve(TOL, 50_000)
    for a, b in zip(base, other):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("block,block_resident,split", [(256, 512, 8 * 1024), (512, 512, 32 * 1024)])
def test_persistent_workgroup_sizes_and_split_solve_to_tolerance(gpu, block, block_resident, split):
    from sovereign_solver_amd.lp.batch import results_of
    probs = uneven_batch(48, seed=9)
    eng = BatchEngine(PackedBatch(probs))
    eng.configure(block=block, block_resident=block_resident, split=split, lines=4)
    out, xs, ys = eng.solve(TOL, 100_000)
    _all_within(probs, results_of(eng, out, xs, ys))


def test_family_cadence_tuned_solve(gpu, cache):
    fam = DispatchFamily(G=20, T=8)
    probs = [fam.problem(k) for k in range(128)]
    res = solve_lp_pdhg_batch(probs, tol=TOL, engine="family", tune=True)
    _all_within(probs[:32], res[:32])
    rec = next(v for k, v in T._load_cache().items() if "/family/" in k)
    assert rec["check_every"] in (25, 50, 100)
    assert all(v["ok"] for v in rec["trials"].values())


def test_tuned_cadence_small_random_lps(gpu, cache):
    probs = [random_sparse_lp(80, 200, 0.05, seed=s) for s in range(3)]
    for p in probs:
        r = solve_lp_pdhg(p, tol=TOL, precision="fp32", tune="auto")
        _all_within([p], [r])
