"""GPU parity of the shared-matrix (family) engine (needs an MI355X).

Every LP must stop at the requested relative KKT as measured on the host by
the reference formula (oracle yardstick), agree with HiGHS (FP64 simplex) on
the objective, and run the same algorithm as the persistent batched engine
(iteration counts within a few checks: the summation orders differ).
"""

import numpy as np
import pytest
from scipy.optimize import linprog

from oracle import yardstick
from oracle.farkas import farkas_value as _farkas_value
from sovereign_solver_amd import solve_lp_pdhg_batch
from sovereign_solver_amd.engine.family import FamilyBatch, FamilyEngine
from sovereign_solver_amd.workloads import DispatchFamily, SetCoverNodes
from tests.parity import assert_objective_parity

pytestmark = pytest.mark.gpu


def _highs(p):
    r = linprog(p.c, A_ub=p.A_ub if p.n_ineq else None, b_ub=p.b_ub if p.n_ineq else None,
                A_eq=p.A_eq if p.n_eq else None, b_eq=p.b_eq if p.n_eq else None,
                bounds=list(zip(p.lb, [None if np.isinf(u) else u for u in p.ub])), method="highs-ds")
    assert r.status == 0
    return r.fun


def _check(probs, res, tol=1e-4, n_highs=6):
    for p, r in zip(probs, res):
        assert r.status == "optimal", (p.name, r.status, r.iterations)
        err = yardstick.rel_kkt(p, r.x, r.y)[0]
        assert err <= 1.02 * tol, (p.name, err, r.rel_kkt)
        assert np.all(r.x >= p.lb - 1e-9) and np.all(r.x <= p.ub + 1e-9)
    for p, r in list(zip(probs, res))[:n_highs]:
        assert_objective_parity(p, r.objective, tol, ref=_highs(p), x=r.x, y=r.y)


@pytest.mark.parametrize("scheme", ["adaptive", "reference"])
def test_family_dispatch_small(gpu, scheme):
    fam = DispatchFamily(G=20, T=6)
    probs = [fam.problem(k) for k in range(70)]                  # two tiles, the second partial
    res = solve_lp_pdhg_batch(probs, tol=1e-4, max_iters=100_000, engine="family", scheme=scheme)
    _check(probs, res)


def test_family_setcover_nodes(gpu):
    sc = SetCoverNodes(rows=100, cols=1000, density=0.01)
    probs = [sc.node(k) for k in range(40)]
    res = solve_lp_pdhg_batch(probs, tol=1e-4, max_iters=100_000, engine="family")
    _check(probs, res)


def test_family_agrees_with_persistent_engine(gpu):
    fam = DispatchFamily(G=30, T=8)
    probs = [fam.problem(k) for k in range(20)]
    rf = solve_lp_pdhg_batch(probs, tol=1e-4, engine="family")
    rp = solve_lp_pdhg_batch(probs, tol=1e-4, engine="persistent")
    for k, (f, p) in enumerate(zip(rf, rp)):
        assert f.status == p.status == "optimal"
        ref = _highs(probs[k])
        assert_objective_parity(probs[k], f.objective, ref=ref, x=f.x, y=f.y)
        assert_objective_parity(probs[k], p.objective, ref=ref, x=p.x, y=p.y)
        assert abs(f.iterations - p.iterations) <= max(200, 0.25 * p.iterations), (f.iterations, p.iterations)


def test_family_launch_path_and_compaction_agree_bitwise(gpu):
    # no atomics and per-lane arithmetic only: neither the launch path (graphs or
    # direct launches) nor moving LPs between tiles changes any LP's result
    fam = DispatchFamily(G=12, T=5)
    fb = FamilyBatch([fam.problem(k) for k in range(150)])
    runs = []
    for graphs, compact in ((False, False), (True, False), (False, True), (True, True)):
        eng = FamilyEngine(fb, use_graphs=graphs, compaction=compact)
        runs.append(eng.solve(1e-4, 100_000))
        if compact:
            assert eng.compactions >= 1
            runs.append(eng.solve(1e-4, 100_000))       # a second solve on the same engine
    a = runs[0]
    for b in runs[1:]:
        assert np.array_equal(a[0]["iterations"], b[0]["iterations"])
        assert np.array_equal(a[0]["status"], b[0]["status"])
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_family_iteration_limit_tail(gpu):
    fam = DispatchFamily(G=12, T=5)
    probs = [fam.problem(k) for k in range(5)]
    res = solve_lp_pdhg_batch(probs, tol=1e-12, max_iters=137, engine="family")
    for p, r in zip(probs, res):
        assert r.status == "iteration_limit" and r.iterations == 137
        assert np.all(np.isfinite(r.x)) and np.all(r.x >= p.lb) and np.all(r.x <= p.ub)


def test_family_warm_start_from_solution(gpu):
    sc = SetCoverNodes(rows=80, cols=600, density=0.015)
    probs = [sc.node(k) for k in range(16)]
    cold = solve_lp_pdhg_batch(probs, tol=1e-4, engine="family")
    warm = solve_lp_pdhg_batch(probs, warm_start=[(r.x, r.y, r.step_scale) for r in cold], tol=1e-4,
                               engine="family")
    for c, w in zip(cold, warm):
        assert w.status == "optimal" and w.iterations <= max(50, c.iterations // 2)


def test_family_device_start_matches_host_start(gpu):
    """pdlp_family_set_start (row-major FP32 source, affine de-normalisation,
    projections, tiling) gives the same solve as the host-side start."""
    import torch
    fam = DispatchFamily(G=12, T=5)
    fb = FamilyBatch([fam.problem(k) for k in range(70)])
    rng = np.random.default_rng(0)
    n, m = fb.n, fb.m
    raw = rng.standard_normal((70, n + m + 1)).astype(np.float32)
    scale = rng.uniform(0.5, 2.0, n + m + 1).astype(np.float32)
    shift = rng.standard_normal(n + m + 1).astype(np.float32)
    val = raw.astype(np.float64) * scale + shift
    val32 = (raw * scale + shift).astype(np.float64)       # the kernel's FP32 arithmetic
    # failover off: this compares the two ways of setting the same start
    a_eng = FamilyEngine(fb, failover_iters=0)
    a_eng.start(1e-4, val32[:, :n], val32[:, n:n + m], np.exp(np.clip(val32[:, -1], -8, 8)))
    a_eng.run(100_000)
    a = a_eng.results()
    b_eng = FamilyEngine(fb, failover_iters=0)
    dv = b_eng.device
    src = torch.from_numpy(raw).to(dv)
    b_eng.start_from_device(1e-4, src, 0, n, n + m, torch.from_numpy(scale).to(dv), torch.from_numpy(shift).to(dv))
    b_eng.run(100_000)
    b = b_eng.results()
    assert np.all(a[0]["status"] == 1) and np.all(b[0]["status"] == 1)
    # the host path rounds x0 / y0 / weight through FP64 first: agree to a few checks
    assert np.all(np.abs(a[0]["iterations"] - b[0]["iterations"]) <= 100)
    assert np.allclose(val, val32, rtol=1e-5, atol=1e-5)


def test_family_garbage_and_nan_device_starts_fail_over(gpu):
    """Network outputs of garbage (large random values, a wild primal weight) and NaN
    through pdlp_family_set_start: the engine compares each start's KKT with the cold
    start's and drops the worse, and restarts cold a start that has made no progress
    by 200 iterations, so every LP still ends optimal within the cold solve's
    iterations plus the failover's 200 (and a check)."""
    import torch
    fam = DispatchFamily(G=20, T=6)
    probs = [fam.problem(k) for k in range(96)]
    fb = FamilyBatch(probs)
    n, m = fb.n, fb.m
    cold = FamilyEngine(fb).solve(1e-4, 100_000)[0]
    assert np.all(cold["status"] == 1)
    rng = np.random.default_rng(7)
    garbage = (100.0 * rng.standard_normal((96, n + m + 1))).astype(np.float32)
    nan = np.full((96, n + m + 1), np.nan, np.float32)
    mixed = np.where(rng.random((96, 1)) < 0.5, garbage, nan).astype(np.float32)
    mixed[:, -1] = 8.0                                   # weight e^8 for every LP
    for src in (garbage, nan, mixed):
        eng = FamilyEngine(fb)
        t = torch.from_numpy(np.ascontiguousarray(src)).to(eng.device)
        eng.start_from_device(1e-4, t, 0, n, n + m)
        eng.run(100_000)
        out, X, Y = eng.results()
        assert np.all(out["status"] == 1)
        assert np.all(out["iterations"] <= cold["iterations"] + 250), (out["iterations"] - cold["iterations"]).max()
        for k in range(0, 96, 12):
            assert yardstick.rel_kkt(probs[k], X[k], Y[k])[0] <= 1.02e-4


def test_family_good_warm_start_survives_the_failover(gpu):
    # solutions as starts: better than the cold start at init, so they are kept
    fam = DispatchFamily(G=20, T=6)
    fb = FamilyBatch([fam.problem(k) for k in range(64)])
    out, X, Y = FamilyEngine(fb).solve(1e-4, 100_000)
    X, Y = X.copy(), Y.copy()
    warm = FamilyEngine(fb).solve(1e-4, 100_000, X, Y, out["weight"])[0]
    assert np.all(warm["status"] == 1)
    # FP32 iterates restarted at the FP64 solution need a few checks for some LPs; none is slower
    # than its cold solve and the mean is a fraction of the cold one
    assert np.all(warm["iterations"] <= out["iterations"])
    assert warm["iterations"].mean() <= 0.35 * out["iterations"].mean()


def test_family_xcd_affinity_and_long_split_agree_bitwise(gpu):
    """Placement of workgroups (XCD affinity) changes no arithmetic; splitting
    rows over a workgroup changes only their summation order."""
    sc = SetCoverNodes(rows=100, cols=1000, density=0.03)
    probs = [sc.node(k) for k in range(150)]                # 3 tiles (fewer than 8 XCDs)
    fb = FamilyBatch(probs, long_threshold=16)
    assert fb.long_rows > 0
    a = FamilyEngine(fb, xcd_affinity=False).solve(1e-4, 100_000)
    b = FamilyEngine(fb, xcd_affinity=True).solve(1e-4, 100_000)
    assert np.array_equal(a[0]["iterations"], b[0]["iterations"])
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    fam = DispatchFamily(G=12, T=5)
    fb2 = FamilyBatch([fam.problem(k) for k in range(600)])   # 10 tiles (more than 8 XCDs)
    c = FamilyEngine(fb2, xcd_affinity=False).solve(1e-4, 100_000)
    d = FamilyEngine(fb2, xcd_affinity=True).solve(1e-4, 100_000)
    assert np.array_equal(c[1], d[1]) and np.array_equal(c[0]["iterations"], d[0]["iterations"])
    r = solve_lp_pdhg_batch(probs, tol=1e-4, engine="family")
    _check(probs, r, n_highs=3)


def test_family_detects_infeasible_nodes_with_certificates(gpu):
    sc = SetCoverNodes(rows=100, cols=1000, density=0.01)
    probs, bad = [], set()
    for k in range(70):
        p = sc.node(k)
        if k % 7 == 3:                               # uncover row k: fix all its variables to 0
            cols = sc.A[k].indices
            lb, ub = p.lb.copy(), p.ub.copy()
            lb[cols] = 0.0
            ub[cols] = 0.0
            p = sc.node_from(lb, ub, name=f"infeasible_{k}")
            bad.add(k)
        probs.append(p)
    res = solve_lp_pdhg_batch(probs, tol=1e-4, max_iters=100_000, engine="family")
    for k, (p, r) in enumerate(zip(probs, res)):
        if k in bad:
            assert r.status == "infeasible", (k, r.status, r.iterations)
            assert _farkas_value(p, r.y) > 0.0
        else:
            assert r.status == "optimal", (k, r.status)
    assert max(r.iterations for k, r in enumerate(res) if k in bad) < 100_000


def _variants(base, count, seed, vary_b=True, vary_c=True):
    """LPs sharing base's matrix objects, with perturbed right-hand sides / costs
    (feasible: the right-hand side stays A x0-shaped for equality families)."""
    from sovereign_solver_amd.problem import Problem
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        b_ub = base.b_ub * (1 + (rng.uniform(-0.1, 0.1, base.b_ub.size) if vary_b else 0))
        b_eq = base.b_eq
        c = base.c * (1 + (rng.uniform(-0.2, 0.2, base.c.size) if vary_c else 0))
        out.append(Problem(c=c, A_ub=base.A_ub, b_ub=b_ub, A_eq=base.A_eq, b_eq=b_eq, lb=base.lb, ub=base.ub,
                           integer_mask=base.integer_mask, name=f"{base.name}_v{k}"))
    return out


def test_family_equality_rows_and_finite_boxes(gpu):
    from sovereign_solver_amd.workloads import random_sparse_lp
    base = random_sparse_lp(80, 200, 0.05, seed=7)          # '=' rows, 0 <= x <= u
    probs = _variants(base, 66, seed=1, vary_b=False)
    res = solve_lp_pdhg_batch(probs, tol=1e-4, max_iters=100_000, engine="family")
    _check(probs, res, n_highs=4)


def test_family_mixed_rows_and_infinite_upper_bounds(gpu):
    from sovereign_solver_amd.problem import Problem
    from sovereign_solver_amd.workloads import transportation_lp
    base = transportation_lp(12, 20, seed=3)                 # supply '<=' rows, demand '=' rows, x >= 0
    rng = np.random.default_rng(2)
    probs = []
    for k in range(40):                                       # more supply keeps every instance feasible
        probs.append(Problem(c=base.c * (1 + rng.uniform(-0.3, 0.3, base.c.size)), A_ub=base.A_ub,
                             b_ub=base.b_ub * (1 + rng.uniform(0.0, 0.3, base.b_ub.size)), A_eq=base.A_eq,
                             b_eq=base.b_eq, lb=base.lb, ub=base.ub, integer_mask=base.integer_mask,
                             name=f"transport_v{k}"))
    res = solve_lp_pdhg_batch(probs, tol=1e-4, max_iters=200_000, engine="family")
    _check(probs, res, n_highs=4)
    fb = FamilyBatch(probs)
    assert fb.is_shared("ub_span") and not fb.is_shared("b")


def _device_rows(eng, name, ln):
    """[B, L] of a per-LP (tiled) or shared device vector, in the problems' own order."""
    from sovereign_solver_amd.engine.family import from_tiles
    fb = eng.fb
    a = eng.t[name].cpu().numpy().astype(np.float64)
    a = np.broadcast_to(a, (fb.count, ln)) if fb.is_shared(name) else from_tiles(a, fb.tiles, ln, fb.count)
    return fb.from_device_cols(np.array(a)) if ln == fb.n else fb.from_device_rows(np.array(a))


def test_family_device_load_matches_the_box_form(gpu):
    """pdlp_family_load tiles the problems' vectors and forms the lifted-form data and
    normalisers: per-LP bounds (set-cover nodes, shared c and b) and per-LP c and b
    (dispatch, shared bounds) against each LP's own box form."""
    from sovereign_solver_amd.engine.forms import box_form
    sc = SetCoverNodes(rows=60, cols=300, density=0.03)
    fam = DispatchFamily(G=12, T=5)
    for probs in ([sc.node(k) for k in range(70)], [fam.problem(k) for k in range(70)]):
        eng = FamilyEngine(FamilyBatch(probs))
        n, m = eng.fb.n, eng.fb.m
        c64, b64, bs = _device_rows(eng, "c64", n), _device_rows(eng, "b64", m), _device_rows(eng, "b_shift", m)
        span, c32 = _device_rows(eng, "ub_span", n), _device_rows(eng, "c", n)
        for k in (0, 33, 69):
            f = box_form(probs[k])
            assert eng.norm_b[k] == pytest.approx(f.norm_b, rel=1e-13)
            assert eng.norm_c[k] == pytest.approx(f.norm_c, rel=1e-13)
            np.testing.assert_array_equal(c64[k], f.c)
            np.testing.assert_array_equal(c32[k], f.c.astype(np.float32))
            np.testing.assert_array_equal(b64[k], f.b)
            np.testing.assert_allclose(bs[k], f.b_shift, rtol=0, atol=1e-12)
            np.testing.assert_array_equal(span[k], f.ub_span if f.ub_span is not None else np.full(n, np.inf))


def test_family_reload_matches_a_fresh_engine(gpu):
    # a resident engine loaded with other LPs of the family solves them as a new engine does
    fam = DispatchFamily(G=12, T=5)
    a, b = [fam.problem(k) for k in range(96)], [fam.problem(1000 + k) for k in range(96)]
    eng = FamilyEngine(FamilyBatch(a))
    eng.solve(1e-4, 100_000)                               # compacts on the way
    eng.load(b)
    o1, x1, y1 = eng.solve(1e-4, 100_000)
    o1, x1, y1 = o1.copy(), x1.copy(), y1.copy()
    o2, x2, y2 = FamilyEngine(FamilyBatch(b)).solve(1e-4, 100_000)
    assert np.all(o1["status"] == 1)
    assert np.array_equal(o1, o2) and np.array_equal(x1, x2) and np.array_equal(y1, y2)


def test_rows_per_wavefront_same_bits():
    """Rows / columns per wavefront change where each row and column runs, not its
    arithmetic: every choice gives the same bits."""
    fam = DispatchFamily(G=40, T=12)
    probs = [fam.problem(10**6 + k) for k in range(70)]
    res = []
    for ipw in (1, 2, 3, 4):
        eng = FamilyEngine(FamilyBatch(probs), check_every=50, items_per_wave=ipw)
        o, X, Y = eng.solve(1e-4, 100_000)
        res.append((o.copy(), X.copy(), Y.copy()))
    o0, X0, Y0 = res[0]
    assert np.all(o0["status"] == 1)
    for o, X, Y in res[1:]:
        assert np.array_equal(o["iterations"], o0["iterations"])
        assert np.array_equal(X, X0) and np.array_equal(Y, Y0)
    assert yardstick.rel_kkt(probs[0], X0[0], Y0[0])[0] <= 1.02e-4
