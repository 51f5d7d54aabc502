"""The serving loop over consecutive batches (needs an MI355X): batch k+1 is
prepared on a worker thread (inputs, staging, upload, network starts) while
batch k solves, on the other of two engines.  Overlap changes no arithmetic,
so every batch gives the bits of the same batch solved alone."""

import numpy as np
import pytest

from oracle import yardstick
from sovereign_solver_amd.learned import nodes as N
from sovereign_solver_amd.learned import warmstart as W
from sovereign_solver_amd.lp.serve import FamilyServer
from sovereign_solver_amd.workloads import DispatchFamily

pytestmark = pytest.mark.gpu


def _batches(fam, count, size):
    return [[fam.problem(10**6 + size * j + k) for k in range(size)] for j in range(count)]


def _check(fam_batches, stream, alone):
    for probs, (o, X, Y, obj), (o1, X1, Y1, obj1) in zip(fam_batches, stream, alone):
        assert np.all(o["status"] == 1)
        assert np.array_equal(o["iterations"], o1["iterations"])
        assert np.array_equal(X, X1) and np.array_equal(Y, Y1) and np.array_equal(obj, obj1)
        for k in (0, len(probs) - 1):
            assert yardstick.rel_kkt(probs[k], X[k], Y[k])[0] <= 1.02e-4
            assert abs(obj[k] - float(probs[k].c @ X[k])) <= 1e-9 * (1 + abs(obj[k]))


def test_serving_loop_cold_same_bits_as_alone(gpu):
    fam = DispatchFamily(G=40, T=12)
    bs = _batches(fam, 3, 64)
    srv = FamilyServer(bs[0], 1e-4, 100_000, engine_kw={"check_every": 50})
    # each item copied before the next is asked for (the views are reused)
    stream = [(o.copy(), X.copy(), Y.copy(), obj.copy()) for o, X, Y, obj in srv.solve_stream(bs)]
    assert len(stream) == 3
    alone = []
    for probs in bs:
        o, X, Y, obj = srv.solve(probs)
        alone.append((o.copy(), X.copy(), Y.copy(), obj.copy()))
    _check(bs, stream, alone)
    assert list(srv.solve_stream([])) == []


@pytest.fixture(scope="module")
def small_model():
    fam = DispatchFamily(G=40, T=12)
    train_ids = np.arange(1024)
    lab = W.label(fam, train_ids)
    return fam, N.train_node_model(fam, lab, stages=2, steps=400, rows=8192, calib_ids=train_ids[-32:])


@pytest.mark.parametrize("where", ["host", "device"])
def test_serving_loop_learned_same_bits_as_alone(gpu, small_model, where):
    """Inputs built on the host, or on the device from the engine's staged vectors:
    the streamed batches give the bits of the same batches solved alone, and both
    input paths give the same starts (the device inputs are bitwise the host's)."""
    fam, model = small_model
    feats = (lambda ps, e: fam.node_inputs_from(ps)) if where == "host" else fam.node_inputs_staged
    bs = _batches(fam, 3, 64)
    srv = FamilyServer(bs[0], 1e-4, 100_000, model=model, features=feats, engine_kw={"check_every": 50})
    stream = [(o.copy(), X.copy(), Y.copy(), obj.copy()) for o, X, Y, obj in srv.solve_stream(bs)]
    alone = []
    for probs in bs:
        o, X, Y, obj = srv.solve(probs)
        alone.append((o.copy(), X.copy(), Y.copy(), obj.copy()))
    _check(bs, stream, alone)
    if where == "device":
        host = FamilyServer(bs[0], 1e-4, 100_000, model=model, features=lambda ps, e: fam.node_inputs_from(ps),
                            engine_kw={"check_every": 50})
        o, X, Y, obj = host.solve(bs[1])
        assert np.array_equal(X, alone[1][1]) and np.array_equal(o["iterations"], alone[1][0]["iterations"])


@pytest.mark.parametrize("learned", [False, True])
def test_streamed_batches_same_bits_as_alone(gpu, small_model, learned):
    """All batches through one engine as one stream (FamilyServer.solve_streamed:
    uploads and network starts on a worker thread, LPs entering tiles as earlier
    ones finish) give each batch's bits when solved alone."""
    fam, model = small_model
    bs = _batches(fam, 3, 64)
    kw = {"model": model, "features": fam.node_inputs_staged} if learned else {}
    srv = FamilyServer(bs[0], 1e-4, 100_000, engine_kw={"check_every": 50}, **kw)
    streamed = srv.solve_streamed(bs)
    assert len(streamed) == 3
    alone = []
    for probs in bs:
        o, X, Y, obj = srv.solve(probs)
        alone.append((o.copy(), X.copy(), Y.copy(), obj.copy()))
    _check(bs, streamed, alone)
    assert srv.solve_streamed([]) == []


@pytest.mark.parametrize("G,T", [(40, 12), (200, 24)])
def test_device_node_inputs_bitwise_host(gpu, G, T):
    """The dispatch family's network inputs computed on the device from the staged
    LP vectors equal the host definition bit for bit, ties in the merit order
    included (equal costs keep the generators' index order)."""
    import torch
    from sovereign_solver_amd.engine.family import FamilyBatch, FamilyEngine
    fam = DispatchFamily(G=G, T=T)
    probs = [fam.problem(10**6 + k) for k in range(70)]
    rng = np.random.default_rng(5)
    for k in range(3):                       # instances with tied costs and demands at capacity steps
        d, c = fam.params(10**7 + k)
        c[rng.choice(G, G // 2, replace=False)] = 33.25
        order = np.argsort(c, kind="stable")
        d[: T // 2] = np.cumsum(fam.pmax[order])[rng.integers(0, G, T // 2)]
        probs.append(fam.problem_from(d, c))
    eng = FamilyEngine(FamilyBatch(probs))
    eng.load(probs)
    V = fam.node_inputs_staged(probs, eng)
    eng.stream.synchronize()
    assert isinstance(V, torch.Tensor)
    ref = fam.node_inputs_from(probs)
    got = V.cpu().numpy()
    assert got.shape == ref.shape and np.array_equal(got.view(np.uint32), ref.view(np.uint32))
