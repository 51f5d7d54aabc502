"""GPU: every config at the size the benchmark runs it (needs an MI355X).

Each result is judged on the host: the reference KKT formula (oracle yardstick, the
explicit reference-form lift, FP64) at the requested tolerance for every LP, and the
objective against the FP64 optimum (HiGHS) on a seeded sample, held to the bar of
tests/parity.py (1e-4, or parity with the reference PDHG at the same tolerance).

* C1: the transportation LP, 1,000 sources x 2,000 sinks (n = 2,000,000);
* C2: all 4,096 uneven random LPs in one batch (two persistent launches);
* C3: the dispatch family at 200 generators x 24 periods, learned warm start
  (the benchmark's node model, trained on 50,000 instances), >= 1.5x fewer iterations;
* C4: 256 node relaxations of the 1,000 x 10,000 set-covering MILP, parent start.
"""

import numpy as np
import pytest

from oracle import yardstick
from sovereign_solver_amd import solve_lp_pdhg, solve_lp_pdhg_batch
from sovereign_solver_amd.learned import nodes as N
from sovereign_solver_amd.learned import warmstart as W
from sovereign_solver_amd.workloads import DispatchFamily, SetCoverNodes, transportation_lp, uneven_batch
from tests.parity import OBJ_TOL, assert_objective_parity, highs, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-4


def _kkt_all(probs, res, tol=TOL):
    errs = []
    for p, r in zip(probs, res):
        assert r.status == "optimal", (p.name, r.status, r.iterations)
        e = yardstick.rel_kkt(p, r.x, r.y)[0]
        errs.append(e)
        assert e <= 1.02 * tol, (p.name, e, r.rel_kkt)
    return np.array(errs)


def test_c1_transport_full_size(gpu):
    p = transportation_lp(1000, 2000, seed=0)
    r = solve_lp_pdhg(p, tol=TOL, precision="fp32", scheme="reference")
    assert r.status == "optimal" and r.iterations < 100_000
    assert yardstick.rel_kkt(p, r.x, r.y)[0] <= 1.02 * TOL
    assert np.all(r.x >= -1e-12)
    # 7.5e-5 from the optimum: within the bar without the reference-PDHG comparison (whose
    # 16,250 iterations on this LP take minutes on the CPU)
    assert rel_err(r.objective, highs(p)) <= OBJ_TOL


def test_c2_all_4096_uneven_lps(gpu):
    probs = uneven_batch(4096, seed=0)
    res = solve_lp_pdhg_batch(probs, tol=TOL, max_iters=100_000, engine="persistent")
    errs = _kkt_all(probs, res)
    print(f"C2: 4096/4096 optimal, host rel KKT max {errs.max():.3e}")
    # HiGHS's simplex takes minutes on some of the larger random LPs: the seeded sample is
    # drawn from the LPs with m <= 400 (1,400 of the 4,096), where it takes a fraction of a second
    rng = np.random.default_rng(2024)
    small = [k for k, p in enumerate(probs) if p.n_constraints <= 400]
    for k in rng.choice(small, size=64, replace=False):
        assert_objective_parity(probs[k], res[k].objective, name=f"uneven_{k}", x=res[k].x, y=res[k].y)


@pytest.mark.timeout(600)
def test_c3_dispatch_full_size_learned(gpu):
    """The benchmark's learned configuration: 50,000 training instances labelled by
    PDHG, a three-stage node model (3-layer width-512 MLP per stage), the primal
    weight calibrated on 256 training instances; 1,024 held-out LPs solved cold and
    from the network's starts.  The learned start takes at least 1.5x fewer
    iterations on average."""
    fam = DispatchFamily()                               # 200 generators x 24 periods
    assert (fam.m, fam.n) == (9224, 4800)
    train_ids, eval_ids = np.arange(50_000), 10**6 + np.arange(1024)
    lab = W.label(fam, train_ids)
    assert np.all(lab.status == "optimal")
    model = N.train_node_model(fam, lab, stages=3, steps=60_000, rows=32768, calib_ids=train_ids[-256:])
    cold, _ = W.solve_family(fam, eval_ids, TOL, 100_000)
    warm, _ = N.solve_nodes(fam, eval_ids, model, TOL, 100_000)
    probs = [fam.problem(int(k)) for k in eval_ids]
    _kkt_all(probs, cold)
    _kkt_all(probs, warm)
    ci = np.array([r.iterations for r in cold])
    wi = np.array([r.iterations for r in warm])
    print(f"C3: held-out iterations cold mean {ci.mean():.1f} max {ci.max()}, learned mean {wi.mean():.1f} "
          f"max {wi.max()}: {ci.mean() / wi.mean():.3f}x")
    assert ci.mean() >= 1.5 * wi.mean()
    for k in range(16):
        ref = highs(probs[k])
        assert_objective_parity(probs[k], warm[k].objective, ref=ref, x=warm[k].x, y=warm[k].y)
        assert_objective_parity(probs[k], cold[k].objective, ref=ref, x=cold[k].x, y=cold[k].y)


def test_c4_setcover_256_nodes_full_size(gpu):
    sc = SetCoverNodes()                                 # 1,000 x 10,000, 0.5% dense
    root = solve_lp_pdhg(sc.root(), tol=TOL, precision="fp32")
    assert root.status == "optimal"
    probs = [sc.node(k) for k in range(256)]
    res = solve_lp_pdhg_batch(probs, warm_start=[(np.clip(root.x, p.lb, p.ub), root.y, root.step_scale)
                                                 for p in probs], tol=TOL, max_iters=100_000, engine="family")
    errs = _kkt_all(probs, res)
    print(f"C4: 256/256 optimal, host rel KKT max {errs.max():.3e}, iterations mean "
          f"{np.mean([r.iterations for r in res]):.0f}")
    for k in range(16):
        assert_objective_parity(probs[k], res[k].objective, x=res[k].x, y=res[k].y)
