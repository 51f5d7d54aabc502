"""A stream of LP batches through one family engine (FamilyEngine.run_stream):
LPs enter free tiles as others finish, and every LP's result is bit for bit
the one a batch solve from the same start gives (its iterations, checks and
decisions are its own; tiles and slots change only where it runs)."""

import numpy as np
import pytest
import torch

from sovereign_solver_amd.engine.batch import OUT_DTYPE
from sovereign_solver_amd.engine.family import FamilyBatch, FamilyEngine, StreamBatch
from sovereign_solver_amd.workloads import DispatchFamily, SetCoverNodes

pytestmark = pytest.mark.gpu


def _sources(eng, probs, dev):
    """LP-major FP64 device sources of the engine's per-LP vectors (b = [b_ub, b_eq])."""
    m_ub = eng.fb.m_ub
    out = {}
    for name in ("c", "b", "lo", "hi"):
        if eng.staged(name) is None:
            continue
        if name == "c":
            a = np.stack([p.c for p in probs])
        elif name == "b":
            a = np.stack([np.concatenate([np.asarray(p.b_ub, float).ravel()[:m_ub],
                                          np.asarray(p.b_eq, float).ravel()]) for p in probs])
        elif name == "lo":
            a = np.stack([p.lb for p in probs])
        else:
            a = np.stack([p.ub for p in probs])
        out[name] = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    return out


def _host(O, X, Y):
    torch.cuda.synchronize()
    o = np.frombuffer(O.cpu().numpy().tobytes(), dtype=OUT_DTYPE).copy()
    return o, X.cpu().numpy(), Y.cpu().numpy()


def _batch(probs, start=None, n=0, max_iters=100_000, tol=1e-4, ce=50):
    eng = FamilyEngine(FamilyBatch(probs), check_every=ce)
    if start is None:
        eng.start(tol)
    else:
        eng.start_from_device(tol, start, x_off=0, y_off=n)
    eng.run(max_iters)
    o, X, Y = eng.results()
    return o.copy(), X.copy(), Y.copy()


def _same(a, b, lo, hi):
    o, X, Y = a
    assert np.array_equal(o[lo:hi], b[0]), "out records differ"
    assert np.array_equal(X[lo:hi], b[1]) and np.array_equal(Y[lo:hi], b[2])


@pytest.mark.parametrize("admit_tiles", [1, 2])
def test_stream_cold_same_bits_as_batches(gpu, admit_tiles):
    fam = DispatchFamily(G=12, T=5)
    sets = [[fam.problem(k) for k in range(128)], [fam.problem(500 + k) for k in range(100)],
            [fam.problem(900 + k) for k in range(64)], [fam.problem(1300 + k) for k in range(30)]]
    eng = FamilyEngine(FamilyBatch(sets[0]), check_every=50)
    got = _host(*eng.run_stream(1e-4, 100_000, [StreamBatch(_sources(eng, s, gpu)) for s in sets],
                                admit_tiles=admit_tiles))
    assert np.all(got[0]["status"] == 1)
    lo = 0
    for s in sets:
        _same(got, _batch(s), lo, lo + len(s))
        lo += len(s)
    # the engine's own batch afterwards: as a fresh engine
    o, X, Y = eng.solve(1e-4, 100_000)
    ref = _batch(sets[0])
    assert np.array_equal(o, ref[0]) and np.array_equal(X, ref[1]) and np.array_equal(Y, ref[2])


def test_stream_engine_with_a_part_tile_same_bits(gpu):
    """An engine of 100 slots (its last tile part empty) streams batches of other
    sizes: admissions never reach past its slots, and every LP keeps its bits."""
    fam = DispatchFamily(G=12, T=5)
    sets = [[fam.problem(2000 + k) for k in range(100)], [fam.problem(2500 + k) for k in range(128)],
            [fam.problem(2900 + k) for k in range(40)]]
    eng = FamilyEngine(FamilyBatch(sets[0]), check_every=50)
    got = _host(*eng.run_stream(1e-4, 100_000, [StreamBatch(_sources(eng, s, gpu)) for s in sets], admit_tiles=1))
    lo = 0
    for s in sets:
        _same(got, _batch(s), lo, lo + len(s))
        lo += len(s)


def test_stream_device_starts_same_bits_as_batches(gpu):
    """Warm starts read on the device per batch (garbage and NaN rows included,
    which the failover judges) give the batch solve's bits."""
    fam = DispatchFamily(G=12, T=5)
    sets = [[fam.problem(k) for k in range(128)], [fam.problem(700 + k) for k in range(128)],
            [fam.problem(1100 + k) for k in range(70)]]
    n, m = fam.n, fam.m
    rng = np.random.default_rng(5)
    starts = []
    for s in sets:
        a = rng.standard_normal((len(s), n + m)).astype(np.float32) * 0.1
        a[3] = 1e6
        a[5, :7] = np.nan
        starts.append(torch.from_numpy(a).to(gpu))
    eng = FamilyEngine(FamilyBatch(sets[0]), check_every=50)
    batches = [StreamBatch(_sources(eng, s, gpu), start=st, x_off=0, y_off=n) for s, st in zip(sets, starts)]
    got = _host(*eng.run_stream(1e-4, 100_000, batches, admit_tiles=1))
    lo = 0
    for s, st in zip(sets, starts):
        _same(got, _batch(s, start=st, n=n), lo, lo + len(s))
        lo += len(s)


def test_stream_iteration_limit_same_bits_as_batches(gpu):
    """Each LP's own limit (PdlpFamily.max_iters) stops it where a batch run to
    that limit does, restarts at the last check included."""
    sc = SetCoverNodes(rows=60, cols=300, density=0.03)
    sets = [[sc.node(k) for k in range(64)], [sc.node(100 + k) for k in range(64)]]
    eng = FamilyEngine(FamilyBatch(sets[0]), check_every=50)
    got = _host(*eng.run_stream(1e-12, 300, [StreamBatch(_sources(eng, s, gpu)) for s in sets], admit_tiles=1))
    lim = got[0]["status"] == 0                        # PDLP_RUNNING: stopped at the limit
    assert lim.sum() >= 64 and np.all(got[0]["iterations"][lim] == 300)   # others: infeasible nodes
    lo = 0
    for s in sets:
        _same(got, _batch(s, tol=1e-12, max_iters=300), lo, lo + len(s))
        lo += len(s)


def test_stream_rejects_bad_batches(gpu):
    fam = DispatchFamily(G=12, T=5)
    probs = [fam.problem(k) for k in range(64)]
    eng = FamilyEngine(FamilyBatch(probs), check_every=50)
    src = _sources(eng, probs, gpu)
    with pytest.raises(ValueError, match="multiple of check_every"):
        eng.run_stream(1e-4, 125, [StreamBatch(src)])
    with pytest.raises(ValueError, match="sources must be exactly"):
        eng.run_stream(1e-4, 100, [StreamBatch({"c": src["c"]})])
    with pytest.raises(ValueError, match="same LPs"):
        StreamBatch(src, count=3)
