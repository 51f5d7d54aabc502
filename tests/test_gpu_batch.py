"""GPU parity of the batched persistent engine (needs an MI355X).

Every LP of a batch must stop at the requested relative KKT as measured on
the host by the reference formula (oracle yardstick), and agree with HiGHS
(FP64 simplex) and with the single-instance engine on the objective.
"""

import numpy as np
import pytest
from scipy.optimize import linprog

from oracle import yardstick
from sovereign_solver_amd import solve_lp_pdhg, solve_lp_pdhg_batch
from sovereign_solver_amd.problem import Problem
from sovereign_solver_amd.workloads import random_sparse_lp, transportation_lp, uneven_batch, uneven_sizes
from tests.parity import assert_objective_parity

pytestmark = pytest.mark.gpu


def _highs(p):
    r = linprog(p.c, A_ub=p.A_ub if p.n_ineq else None, b_ub=p.b_ub if p.n_ineq else None,
                A_eq=p.A_eq if p.n_eq else None, b_eq=p.b_eq if p.n_eq else None,
                bounds=list(zip(p.lb, [None if np.isinf(u) else u for u in p.ub])), method="highs-ds")
    assert r.status == 0
    return r.fun


@pytest.mark.parametrize("block", [256, 512])
def test_batch_uneven_small(gpu, block):
    probs = uneven_batch(24, m_lo=20, m_hi=300, density=0.03, seed=5)
    res = solve_lp_pdhg_batch(probs, tol=1e-4, max_iters=100_000, block=block)
    for p, r in zip(probs, res):
        assert r.status == "optimal", (p.name, r.status, r.iterations)
        err = yardstick.rel_kkt(p, r.x, r.y)[0]
        assert err <= 1.02e-4, (p.name, err, r.rel_kkt)
        assert_objective_parity(p, r.objective, ref=_highs(p), x=r.x, y=r.y)


def test_batch_mixed_row_kinds_and_bounds(gpu):
    probs = [transportation_lp(8, 12, seed=s) for s in range(3)] + [random_sparse_lp(30, 70, 0.1, seed=s)
                                                                   for s in range(3)]
    res = solve_lp_pdhg_batch(probs, tol=1e-4)
    for p, r in zip(probs, res):
        assert r.status == "optimal"
        assert yardstick.rel_kkt(p, r.x, r.y)[0] <= 1.02e-4
        assert np.all(r.x >= p.lb - 1e-9) and np.all(r.x <= p.ub + 1e-9)
        assert_objective_parity(p, r.objective, ref=_highs(p), x=r.x, y=r.y)


def test_batch_agrees_with_single_engine(gpu):
    probs = [random_sparse_lp(100, 250, 0.03, seed=s) for s in range(4)]
    rb = solve_lp_pdhg_batch(probs, tol=1e-4)
    for p, b in zip(probs, rb):
        s = solve_lp_pdhg(p, tol=1e-4, precision="fp32")
        assert s.status == b.status == "optimal"
        ref = _highs(p)
        assert_objective_parity(p, s.objective, ref=ref, x=s.x, y=s.y)
        assert_objective_parity(p, b.objective, ref=ref, x=b.x, y=b.y)
        # same algorithm: iteration counts agree to a few checks (summation order differs)
        assert abs(s.iterations - b.iterations) <= max(200, 0.2 * s.iterations)


def test_batch_warm_start_from_solution(gpu):
    probs = [random_sparse_lp(80, 200, 0.04, seed=s) for s in range(5)]
    cold = solve_lp_pdhg_batch(probs, tol=1e-4)
    warm = solve_lp_pdhg_batch(probs, warm_start=[(r.x, r.y, r.step_scale) for r in cold], tol=1e-4)
    for c, w in zip(cold, warm):
        # the cold result sits just under the tolerance and FP32 iterates restart near it,
        # so the warm solve needs a few checks, still far fewer iterations than the cold one
        assert w.status == "optimal" and w.iterations <= c.iterations // 2


def test_batch_iteration_limit_and_trivial(gpu):
    probs = [random_sparse_lp(60, 150, 0.05, seed=1), Problem.empty(4), random_sparse_lp(40, 100, 0.05, seed=2)]
    res = solve_lp_pdhg_batch(probs, tol=1e-9, max_iters=120)
    assert res[1].status == "optimal" and res[1].iterations == 0 and np.array_equal(res[1].x, np.zeros(4))
    for r in (res[0], res[2]):
        assert r.status == "iteration_limit" and r.iterations == 120
        assert np.all(np.isfinite(r.x))


def test_batch_larger_lp_in_lds_limit(gpu):
    # the largest size of the uneven family (m = 2000, n = 5000) in one workgroup
    p = random_sparse_lp(2000, 5000, 0.01, seed=3, family=2)
    r = solve_lp_pdhg_batch([p], tol=1e-4, max_iters=100_000)[0]
    assert r.status == "optimal"
    assert yardstick.rel_kkt(p, r.x, r.y)[0] <= 1.02e-4


def test_batch_is_deterministic_and_parts_compose(gpu):
    from sovereign_solver_amd.engine.batch import BatchEngine, PackedBatch
    probs = uneven_batch(48, m_lo=50, m_hi=700, density=0.02, seed=9)
    pk = PackedBatch(probs)
    assert 0 < pk.n_resident < len(probs)                  # both parts present
    eng = BatchEngine(pk)
    o1, x1, y1 = eng.solve(1e-4, 50_000)
    assert np.all(o1["status"] == 1)
    o2, x2, y2 = eng.solve(1e-4, 50_000)
    # gathers in a fixed order, no atomics: the same bits every run
    assert np.array_equal(x1, x2) and np.array_equal(y1, y2) and np.array_equal(o1, o2)
    # the two parts launched one at a time give the same bits as one call
    eng.set_warm_start(None, None, None)
    eng.launch(1e-4, 50_000, parts=1)
    eng.launch(1e-4, 50_000, parts=2)
    o3, x3, y3 = eng.results()
    assert np.array_equal(x1, x3) and np.array_equal(y1, y3) and np.array_equal(o1, o3)


def test_batch_fp64_and_fp32_parts_agree(gpu, monkeypatch):
    import sovereign_solver_amd.engine.batch as B
    probs = [random_sparse_lp(120, 300, 0.03, seed=s) for s in range(6)]
    r64 = solve_lp_pdhg_batch(probs, tol=1e-5, engine="persistent")      # small: all in the FP64 part
    monkeypatch.setattr(B, "RESIDENT_LDS", 0)                             # now all in the FP32 part
    assert B.PackedBatch(probs).n_resident == 0
    r32 = solve_lp_pdhg_batch(probs, tol=1e-5, engine="persistent")
    for p, a, b in zip(probs, r64, r32):
        assert a.status == b.status == "optimal", (a.status, b.status)
        ref = _highs(p)
        for r in (a, b):
            assert yardstick.rel_kkt(p, r.x, r.y)[0] <= 1.02e-5
            assert abs(r.objective - ref) <= 1e-4 * (1 + abs(ref)), (r.objective, ref)


def test_batch_small_lps_with_empty_columns_converge(gpu):
    # three LPs of the uneven 4,096-LP batch (m = 50..64, n = 2.5m, 1% dense: most
    # columns empty) whose FP32 iterates drove the primal weight to its bound
    # without converging in 200,000 iterations; FP64 iterates finish them
    ms = uneven_sizes(4096, 50, 2000, 0)
    probs = [random_sparse_lp(int(ms[k]), int(round(2.5 * ms[k])), 0.01, seed=k, family=2) for k in (655, 2272, 2384)]
    assert [p.n_constraints for p in probs] == [64, 52, 50]
    res = solve_lp_pdhg_batch(probs, tol=1e-4, max_iters=50_000, engine="persistent")
    for p, r in zip(probs, res):
        assert r.status == "optimal" and r.iterations <= 20_000, (r.status, r.iterations)
        assert yardstick.rel_kkt(p, r.x, r.y)[0] <= 1.02e-4
        assert_objective_parity(p, r.objective, ref=_highs(p), x=r.x, y=r.y)


def test_batch_fp32_part_layouts_agree_bitwise(gpu):
    # x, its partial sum, tau_hat and c in LDS or in global memory: same arithmetic, same bits
    from sovereign_solver_amd.engine.batch import BatchEngine, PackedBatch, lds_bytes_of
    probs = uneven_batch(24, m_lo=300, m_hi=900, density=0.02, seed=4)
    pk = PackedBatch(probs)
    assert pk.n_resident < len(probs) and pk.x_in_lds
    o1, x1, y1 = BatchEngine(pk).solve(1e-4, 50_000)
    pk.set_x_in_lds(False)
    assert pk.lds_bytes >= int(lds_bytes_of(pk.ms, pk.ns)[~pk.resident].max())
    o2, x2, y2 = BatchEngine(pk).solve(1e-4, 50_000)
    assert np.all(o1["status"] == 1)
    assert np.array_equal(x1, x2) and np.array_equal(y1, y2) and np.array_equal(o1, o2)


@pytest.mark.parametrize("method", ["ruiz", "ruiz_pc"])
def test_scaled_batch(gpu, method):
    """Scaled persistent batch: every LP stops at the requested KKT of the
    original LP (host yardstick on the unscaled x, y), and three small LPs of
    the uneven batch that the unscaled scheme needs 6,000-15,000 iterations for
    finish in a fraction of that.  Those three are ill-conditioned (86 of the
    170 columns of m = 68 empty, finite bounds): a relative KKT of 7e-5 (primal
    residual 0.006 against a normaliser of 90) leaves the objective 4% from
    HiGHS's — where the reference's own PDHG at the same tolerance lands as far
    (4.0% on LP 3579): each LP is held to the objective bar of tests/parity.py."""
    slow = [(94, 3721), (68, 3579), (68, 1498)]               # (m, index in the 4,096-LP uneven batch)
    probs = [random_sparse_lp(m, round(2.5 * m), 0.01, seed=s, family=2) for m, s in slow]
    probs += [random_sparse_lp(m, round(2.5 * m), 0.01, seed=100 + k, family=2)
              for k, m in enumerate((60, 150, 300, 500, 800))]
    plain = solve_lp_pdhg_batch(probs, tol=1e-4, max_iters=100_000, engine="persistent")
    scaled = solve_lp_pdhg_batch(probs, tol=1e-4, max_iters=100_000, engine="persistent", scaling=method)
    for k, (p, r) in enumerate(zip(probs, scaled)):
        assert r.status == "optimal", (p.name, r.status, r.iterations)
        err = yardstick.rel_kkt(p, r.x, r.y)[0]
        assert err <= 1.05e-4, (p.name, err, r.rel_kkt)
        assert abs(err - r.rel_kkt) <= 1e-3 * err + 1e-9     # the device measured the original LP too
        assert_objective_parity(p, r.objective, ref=_highs(p), x=r.x, y=r.y)
    assert sum(r.iterations for r in scaled[:3]) < 0.3 * sum(r.iterations for r in plain[:3])


def test_device_arena_transpose_bitwise_host(gpu):
    """K^T of every stored matrix built on the device (pdlp_arena_transpose) equals the
    host packer's one-transpose-of-the-union, bit for bit: column pointers, 16-bit
    row indices, FP32 values — with matrices shared between LPs and an LP whose
    matrix has an empty row and an empty column."""
    import scipy.sparse as sp
    from sovereign_solver_amd.engine.batch import BatchEngine, PackedBatch
    from sovereign_solver_amd.problem import Problem
    from sovereign_solver_amd.workloads import uneven_batch
    probs = uneven_batch(48, seed=7)
    probs += [probs[3]] * 3                                       # one matrix stored once for 4 LPs
    A = sp.csr_matrix(np.array([[1.0, 0.0, 2.0], [0.0, 0.0, 0.0], [3.0, 0.0, -1.0]]))
    probs.append(Problem(c=np.ones(3), A_ub=A, b_ub=np.ones(3), A_eq=sp.csr_matrix((0, 3)), b_eq=np.zeros(0),
                         lb=np.zeros(3), ub=np.full(3, np.inf), integer_mask=np.zeros(3, bool)))
    host = PackedBatch(probs, transpose="host")
    eng = BatchEngine(PackedBatch(probs))
    t = eng.t
    assert np.array_equal(t["t_col_ptr"].cpu().numpy(), host.t_col_ptr)
    assert np.array_equal(t["t_row_idx"].cpu().numpy().view(np.uint16), host.t_row_idx)
    assert np.array_equal(t["t_val"].cpu().numpy(), host.t_val.astype(np.float32))
