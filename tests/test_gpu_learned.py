"""What the learned warm start gains (needs an MI355X).

The network is trained on the GPU by the library's own MFMA forward/backward
and Adam kernels on solutions PDHG itself produced; held-out instances are
then solved cold and from the network's prediction by the same batched
engine, which finishes and judges every solve (reference KKT on the host).
"""

import numpy as np
import pytest

from oracle import yardstick
from sovereign_solver_amd.learned import warmstart as W
from sovereign_solver_amd.workloads import DispatchFamily, SetCoverNodes

from tests.parity import assert_objective_parity, highs

pytestmark = pytest.mark.gpu


def _check_all(fam, ids, res, tol=1e-4, k=6):
    for r in res:
        assert r.status == "optimal"
    for kk in range(min(k, len(ids))):
        p = fam.problem(int(ids[kk]))
        assert yardstick.rel_kkt(p, res[kk].x, res[kk].y)[0] <= 1.02 * tol


def test_trained_network_gains_over_cold_start(gpu):
    fam = DispatchFamily(G=60, T=12)
    train_ids, eval_ids = np.arange(8192), 10**6 + np.arange(256)
    lab = W.label(fam, train_ids)
    assert np.all(lab.status == "optimal")
    mlp, nin, nout, losses = W.train(fam.features_of(train_ids), W.targets_of(lab), d_h=256, epochs=30,
                                     batch=256, lr=1e-3, out_blocks=W.output_blocks(fam))
    assert losses[-1] < 0.5 * losses[0], losses
    # the prediction is much closer to the held-out solutions than the cold start
    lab_ev = W.label(fam, eval_ids)
    pred = W.predict(mlp, nin, nout, fam.features_of(eval_ids))
    x_err = np.linalg.norm(np.clip(pred[:, : fam.n], 0, np.tile(fam.pmax, fam.T)) - lab_ev.x)
    assert x_err < 0.5 * np.linalg.norm(lab_ev.x), x_err
    cold, _ = W.solve_family(fam, eval_ids, 1e-4, 100_000)
    warm, _ = W.solve_learned(fam, eval_ids, W.LearnedStart(mlp, nin, nout), 1e-4, 100_000, use_y=False)
    _check_all(fam, eval_ids, cold)
    _check_all(fam, eval_ids, warm)
    ci = np.array([r.iterations for r in cold])
    wi = np.array([r.iterations for r in warm])
    print(f"held-out iterations: cold mean {ci.mean():.0f} max {ci.max()}, learned mean {wi.mean():.0f} max {wi.max()}")
    assert wi.mean() < ci.mean()
    assert wi.max() <= ci.max()          # the slowest LP sets the time of a lockstep batch
    for k in range(6):
        p = fam.problem(int(eval_ids[k]))
        ref = highs(p)
        assert_objective_parity(p, cold[k].objective, ref=ref, x=cold[k].x, y=cold[k].y)
        assert_objective_parity(p, warm[k].objective, ref=ref, x=warm[k].x, y=warm[k].y)


def test_parent_warm_start_gains_on_nodes(gpu):
    from sovereign_solver_amd import solve_lp_pdhg
    sc = SetCoverNodes(rows=200, cols=2000, density=0.005)
    root = solve_lp_pdhg(sc.root(), tol=1e-4, precision="fp32")
    assert root.status == "optimal"
    ids = np.arange(64)
    probs = [sc.node(int(k)) for k in ids]
    cold, _ = W.solve_family(sc, ids, 1e-4, 100_000)
    warm, _ = W.solve_family(sc, ids, 1e-4, 100_000,
                             warm=[(np.clip(root.x, p.lb, p.ub), root.y, root.step_scale) for p in probs])
    _check_all(sc, ids, cold)
    _check_all(sc, ids, warm)
    ci = np.mean([r.iterations for r in cold])
    wi = np.mean([r.iterations for r in warm])
    print(f"nodes: cold mean {ci:.0f}, parent warm start mean {wi:.0f}")
    assert wi < ci
