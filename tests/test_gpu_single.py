"""GPU parity of the single-instance engine (needs an MI355X).

Reference form, FP64, exact arithmetic: the reference's published PDHG rows
(status, iteration count, objective to 8 digits) and the oracle trajectory.
Box form, FP32/FP64: termination at the requested relative KKT measured on
the host by the reference formula, objective against HiGHS (FP64 simplex).
"""

import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import torch
from scipy.optimize import linprog

from oracle import pdhg_ref, yardstick
from sovereign_solver_amd import solve_lp_pdhg, solve_lp_pdhg_gpu, get_available_device
from sovereign_solver_amd.engine import _lib as L
from sovereign_solver_amd.problem import Problem
from sovereign_solver_amd.workloads import random_sparse_lp, transportation_lp
from tests.golden.stress_instances import golden_problem, load_published
from tests.parity import assert_objective_parity

pytestmark = pytest.mark.gpu
PUB = load_published()


def _highs_obj(p):
    r = linprog(p.c, A_ub=p.A_ub if p.n_ineq else None, b_ub=p.b_ub if p.n_ineq else None,
                A_eq=p.A_eq if p.n_eq else None, b_eq=p.b_eq if p.n_eq else None,
                bounds=list(zip(p.lb, [None if np.isinf(u) else u for u in p.ub])), method="highs-ds")
    assert r.status == 0
    return r.fun


# ----------------------------------------------------------------- SpMV kernel
@pytest.mark.parametrize("tpr", [1, 2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_spmv_matches_scipy(gpu, tpr, prec):
    rng = np.random.default_rng(tpr)
    A = sp.random(3000, 1700, density=0.01, random_state=tpr, format="csr")
    A.data = rng.standard_normal(A.nnz)
    A.indices = A.indices.astype(np.int32)
    A.indptr = A.indptr.astype(np.int32)
    x = rng.standard_normal(1700)
    dt = torch.float32 if prec == "fp32" else torch.float64
    d = lambda a, t: torch.from_numpy(np.ascontiguousarray(a)).to(gpu, t)
    rp, ci, v, xv = d(A.indptr, torch.int32), d(A.indices, torch.int32), d(A.data, dt), d(x, dt)
    y = torch.zeros(3000, dtype=dt, device=gpu)
    L.check(L.lib().pdlp_spmv_csr(L.PDLP_F32 if prec == "fp32" else L.PDLP_F64, rp.data_ptr(), ci.data_ptr(),
                                  v.data_ptr(), xv.data_ptr(), y.data_ptr(), 3000, tpr, None), "spmv")
    torch.cuda.synchronize()
    got = y.double().cpu().numpy()
    if prec == "fp64":
        Ad, xd = A, x
        np.testing.assert_allclose(got, Ad @ xd, rtol=0, atol=1e-12 * (abs(A) @ abs(x)).max())
    else:
        A32 = A.copy()
        A32.data = A.data.astype(np.float32).astype(np.float64)
        x32 = x.astype(np.float32).astype(np.float64)
        k = np.diff(A.indptr).max()
        bound = (k + 2) * 2.0**-24 * (abs(A32) @ abs(x32)) + 1e-30
        assert np.all(np.abs(got - A32 @ x32) <= bound)


def test_spmv_f64acc(gpu):
    rng = np.random.default_rng(0)
    A = sp.random(500, 800, density=0.05, random_state=0, format="csr")
    A.data = rng.standard_normal(A.nnz).astype(np.float32).astype(np.float64)
    x = rng.standard_normal(800)
    d = lambda a, t: torch.from_numpy(np.ascontiguousarray(a)).to(gpu, t)
    # every device buffer stays referenced until the kernel has finished: a
    # temporary freed after data_ptr() is recycled by the next upload
    y = torch.zeros(500, dtype=torch.float64, device=gpu)
    rp = d(A.indptr.astype(np.int32), torch.int32)
    ci = d(A.indices.astype(np.int32), torch.int32)
    v = d(A.data, torch.float32)
    xv = d(x, torch.float64)
    assert rp.numel() == 501 and int(rp[-1]) == ci.numel() == v.numel() and xv.numel() == 800
    L.check(L.lib().pdlp_spmv_csr_f64acc(rp.data_ptr(), ci.data_ptr(), v.data_ptr(), xv.data_ptr(),
                                         y.data_ptr(), 500, 8, None), "f64acc")
    torch.cuda.synchronize()
    np.testing.assert_allclose(y.cpu().numpy(), A @ x, rtol=1e-12, atol=1e-12)


# ---------------------------------------------- reference form, exact FP64
@pytest.mark.parametrize("name", ["blending", "afiro", "random_n50_m30", "brandy", "adlittle"])
def test_exact_reference_form_reproduces_published_rows(gpu, name):
    row = PUB[(name, "pdhg")]
    r = solve_lp_pdhg(golden_problem(name), max_iters=int(row["max_iters"]), tol=1e-6,
                      formulation="reference", precision="fp64", exact=True)
    assert r.status == row["status"]
    assert r.iterations == int(row["iterations"])
    pub = float(row["objective"])
    assert abs(r.objective - pub) <= 6e-9 * max(1.0, abs(pub)), (r.objective, pub)


def test_exact_reference_form_sparse_n500(gpu):
    row = PUB[("sparse_n500_m250", "pdhg")]
    r = solve_lp_pdhg(golden_problem("sparse_n500_m250"), formulation="reference", precision="fp64", exact=True)
    assert (r.status, r.iterations) == ("optimal", int(row["iterations"]))
    assert abs(r.objective - float(row["objective"])) <= 6e-9 * abs(float(row["objective"]))


def test_exact_trajectory_matches_oracle_iterates(gpu):
    p = random_sparse_lp(60, 150, 0.05, seed=4)
    for iters in (1, 50, 1234):
        o = pdhg_ref.solve(p, max_iters=iters, tol=1e-12)
        g = solve_lp_pdhg(p, max_iters=iters, tol=1e-12, formulation="reference", precision="fp64", exact=True)
        assert g.status == o.status == "iteration_limit"
        np.testing.assert_allclose(g.x, o.x, rtol=1e-13, atol=1e-13)


# ------------------------------------------------------------- box form
@pytest.mark.parametrize("seed", [0, 1])
def test_box_fp64_random_lp(gpu, seed):
    p = random_sparse_lp(200, 500, 0.02, seed=seed)
    r = solve_lp_pdhg(p, tol=1e-6, precision="fp64")
    assert r.status == "optimal"
    err = yardstick.rel_kkt(p, r.x, r.y)[0]
    assert err <= 1.0001e-6
    ref = _highs_obj(p)
    assert abs(r.objective - ref) <= 1e-4 * (1 + abs(ref))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_box_fp32_random_lp(gpu, seed):
    p = random_sparse_lp(200, 500, 0.02, seed=seed)
    r = solve_lp_pdhg(p, tol=1e-4, precision="fp32")
    assert r.status == "optimal"
    err = yardstick.rel_kkt(p, r.x, r.y)[0]
    assert err <= 1.02e-4, err
    assert abs(r.rel_kkt - err) <= 1e-6 + 1e-3 * err
    assert_objective_parity(p, r.objective, x=r.x, y=r.y)


def test_box_fp32_transport_small(gpu):
    p = transportation_lp(50, 100, seed=0)
    r = solve_lp_pdhg(p, tol=1e-4, precision="fp32", max_iters=200_000)
    assert r.status == "optimal"
    assert yardstick.rel_kkt(p, r.x, r.y)[0] <= 1.02e-4
    assert_objective_parity(p, r.objective, x=r.x, y=r.y)


def test_graphs_and_direct_launches_agree_bitwise(gpu):
    p = random_sparse_lp(100, 250, 0.03, seed=9)
    a = solve_lp_pdhg(p, tol=1e-4, precision="fp32", use_graphs=True)
    b = solve_lp_pdhg(p, tol=1e-4, precision="fp32", use_graphs=False)
    assert a.iterations == b.iterations
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


def test_warm_start_from_solution_stops_at_first_check(gpu):
    p = random_sparse_lp(100, 250, 0.03, seed=3)
    a = solve_lp_pdhg(p, tol=1e-4, precision="fp64")
    b = solve_lp_pdhg(p, tol=1e-4, precision="fp64", x0=a.x, y0=a.y, step_scale=a.step_scale)
    assert b.status == "optimal" and b.iterations <= 100 < a.iterations


def test_max_iters_not_multiple_of_check(gpu):
    p = random_sparse_lp(50, 120, 0.05, seed=1)
    o = pdhg_ref.solve(p, max_iters=137, tol=1e-12)
    g = solve_lp_pdhg(p, max_iters=137, tol=1e-12, formulation="reference", precision="fp64", exact=True)
    assert g.iterations == 137 and g.status == "iteration_limit"
    np.testing.assert_allclose(g.x, o.x, rtol=1e-13, atol=1e-13)


# ------------------------------------------------------- drop-in behaviour
def _make(c, A_ub, b_ub):
    A = np.array(A_ub, float)
    n = A.shape[1]
    return Problem(c=np.array(c, float), A_ub=sp.csr_matrix(A), b_ub=np.array(b_ub, float),
                   A_eq=sp.csr_matrix((0, n)), b_eq=np.zeros(0), lb=np.zeros(n), ub=np.full(n, np.inf),
                   integer_mask=np.zeros(n, bool))


@pytest.mark.parametrize("c,A,b,expected", [
    ([-3, -5], [[1, 0], [0, 2], [3, 2]], [4, 12, 18], -36.0),
    ([1, 1], [[1, 0], [0, 1]], [3, 3], 0.0),
    ([-1, -2], [[1, 1]], [4], -8.0),
])
def test_reference_test_lps(gpu, c, A, b, expected):
    """The LPs of tests/test_pdhg.py:81-99 with tight checks instead of |Δ|<0.5."""
    p = _make(c, A, b)
    r = solve_lp_pdhg_gpu(p, device="auto")
    assert r.status == "optimal"
    assert abs(r.objective - expected) <= 1e-4 * (1 + abs(expected))
    assert np.all(r.x >= -1e-12)


def test_device_resolution(gpu, caplog):
    """tests/test_pdhg_gpu.py:148-161 of the reference: an explicit device the engine
    does not have warns and still returns a result (here computed on the HIP GPU);
    an unknown device string raises ValueError."""
    assert get_available_device() == "cuda"
    p = _make([-1.0, -2.0], [[1.0, 1.0]], [4.0])
    for dev in ("cpu", "mps", "cuda", "auto"):
        with caplog.at_level("WARNING"):
            r = solve_lp_pdhg_gpu(p, device=dev)
        assert r.status == "optimal" and abs(r.objective + 8.0) <= 1e-4 * 9, (dev, r.status, r.objective)
        if dev in ("cpu", "mps"):
            assert "HIP GPU only" in caplog.text
    with pytest.raises(ValueError):
        solve_lp_pdhg_gpu(p, device="tpu")


def test_garbage_and_nan_warm_starts_fail_over(gpu):
    """A warm start worse than the cold start is dropped at once, so a garbage or NaN
    start costs no more than the failover's 200 iterations over the cold solve."""
    p = random_sparse_lp(100, 250, 0.03, seed=3)
    cold = solve_lp_pdhg(p, tol=1e-4, precision="fp32")
    rng = np.random.default_rng(1)
    starts = [(1e3 * rng.standard_normal(p.n_vars), 1e3 * rng.standard_normal(p.n_constraints), 5e3),
              (np.full(p.n_vars, np.nan), np.full(p.n_constraints, np.nan), float("nan")),
              (rng.uniform(0, 1, p.n_vars) * p.ub, None, None)]
    for x0, y0, w in starts:
        r = solve_lp_pdhg(p, tol=1e-4, precision="fp32", x0=x0, y0=y0, step_scale=w)
        assert r.status == "optimal", r.status
        assert yardstick.rel_kkt(p, r.x, r.y)[0] <= 1.02e-4
        assert r.iterations <= cold.iterations + 250, (r.iterations, cold.iterations)


def test_empty_problem_returns_optimal_zeros(gpu):
    r = solve_lp_pdhg(Problem.empty(3))
    assert r.status == "optimal" and r.iterations == 0 and np.array_equal(r.x, np.zeros(3))


def test_sell_dual_step_matches_csr_bitwise(gpu):
    """The SELL-64 dual step sums each row in CSR order, as the one-lane CSR
    kernel does: the whole solve is bitwise the same."""
    from sovereign_solver_amd.engine.forms import box_form
    from sovereign_solver_amd.engine.single import SingleEngine
    from sovereign_solver_amd.engine.tune import Tune
    f = box_form(transportation_lp(50, 100, seed=0))
    base = Tune(tpr_primal=1, tpr_dual=1, block=256, check_every=50)
    a = SingleEngine(f, "fp32", tune=base).solve(1e-4, 50_000)
    eng = SingleEngine(f, "fp32", tune=base.with_(sell=1))
    assert eng.sell_slices >= 1 and eng.n_csr_rows > 0
    b = eng.solve(1e-4, 50_000)
    assert a["status"] == b["status"] == "optimal"
    assert a["iterations"] == b["iterations"]
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["y"], b["y"])
    # the other lane counts of the CSR part agree to rounding
    c = SingleEngine(f, "fp32", tune=base.with_(sell=1, tpr_dual=8)).solve(1e-4, 50_000)
    assert c["status"] == "optimal" and abs(c["iterations"] - a["iterations"]) <= 500


@pytest.mark.parametrize("tc", [64, 256, 1024])
@pytest.mark.parametrize("mode", [2, 3, 4, 5, 6, 7])
def test_workgroup_sell_dual_step_agrees(gpu, mode, tc):
    """Slices and CSR rows in one launch, a workgroup per slice piece (entries split
    among its wavefronts) and tc threads per CSR row (a wavefront, several, or the
    whole workgroup): the same optimum as the CSR step, every run the same bits,
    and the iteration counted once per step."""
    from sovereign_solver_amd.engine.forms import box_form
    from sovereign_solver_amd.engine.single import SingleEngine
    from sovereign_solver_amd.engine.tune import Tune
    p = transportation_lp(60, 200, seed=2)
    f = box_form(p)
    base = Tune(tpr_primal=1, tpr_dual=8, block=256, check_every=50)
    a = SingleEngine(f, "fp32", tune=base).solve(1e-4, 50_000)
    eng = SingleEngine(f, "fp32", tune=base.with_(sell=mode, tpr_dual=tc))
    assert eng.sell_slices >= 1 and eng.n_csr_rows > 0
    b = eng.solve(1e-4, 50_000)
    c = eng.solve(1e-4, 50_000)
    assert a["status"] == b["status"] == "optimal"
    assert abs(a["iterations"] - b["iterations"]) <= 500
    assert b["iterations"] == c["iterations"] and np.array_equal(b["x"], c["x"]) and np.array_equal(b["y"], c["y"])
    assert b["ctrl"].iter == b["iterations"] or b["ctrl"].done_iter == b["iterations"]
    assert yardstick.rel_kkt(p, b["x"], b["y"])[0] <= 1.02e-4
    ref = _highs_obj(p)
    assert_objective_parity(p, float(p.c @ b["x"]), ref=ref, x=b["x"], y=b["y"])


@pytest.mark.parametrize("kc", [8, 32])
def test_split_sell_dual_step_agrees(gpu, kc):
    """The split SELL step (partial sums per chunk of kc entries) agrees with the
    CSR step to rounding and reaches the same optimum."""
    from sovereign_solver_amd.engine.forms import box_form
    from sovereign_solver_amd.engine.single import SingleEngine
    from sovereign_solver_amd.engine.tune import Tune
    p = transportation_lp(50, 100, seed=0)
    f = box_form(p)
    base = Tune(tpr_primal=1, tpr_dual=8, block=256, check_every=50)
    a = SingleEngine(f, "fp32", tune=base).solve(1e-4, 50_000)
    b = SingleEngine(f, "fp32", tune=base.with_(sell=kc)).solve(1e-4, 50_000)
    assert a["status"] == b["status"] == "optimal"
    assert abs(a["iterations"] - b["iterations"]) <= 500
    assert yardstick.rel_kkt(p, b["x"], b["y"])[0] <= 1.02e-4
    ref = _highs_obj(p)
    assert_objective_parity(p, float(p.c @ b["x"]), ref=ref, x=b["x"], y=b["y"])


@pytest.mark.parametrize("lp", ["transport", "random"])
def test_sell_primal_step_same_bits_as_csr(gpu, lp, monkeypatch):
    """The primal step over the SELL-64 copy of K^T (one lane per column, one or two
    slices per wavefront) sums each column in CSR order, so whole solves give the
    bits of the one-lane CSR kernel:
    a transportation LP (every column 2 entries, no padding) and a random LP whose
    unequal columns pad the slices (the padding limit lifted to build it)."""
    from sovereign_solver_amd.engine import sell as SL
    from sovereign_solver_amd.engine.forms import box_form
    from sovereign_solver_amd.engine.single import SingleEngine
    from sovereign_solver_amd.engine.tune import Tune
    monkeypatch.setattr(SL, "COLUMN_PADDING_MAX", 100.0)
    p = transportation_lp(60, 200, seed=5) if lp == "transport" else random_sparse_lp(150, 400, 0.03, seed=7)
    f = box_form(p)
    base = Tune(tpr_primal=1, tpr_dual=8, block=256, check_every=50, scheme=1)
    runs = []
    for sp in (0, 1, 2):                     # CSR, one slice per wavefront, two slices per wavefront
        eng = SingleEngine(f, "fp32", tune=base.with_(sell_primal=sp))
        assert eng.t_sell_slices == -(-f.n // 64)
        r = eng.solve(1e-4, 50_000)
        runs.append(r)
        eng.drop_graphs()
    a, b, c = runs
    assert a["status"] == b["status"] == c["status"] == "optimal"
    assert a["iterations"] == b["iterations"] == c["iterations"]
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["y"], b["y"])
    assert np.array_equal(a["x"], c["x"]) and np.array_equal(a["y"], c["y"])
    e = yardstick.rel_kkt(p, *extract_xy(f, p, b))[0]
    assert e <= 1.02e-4


def extract_xy(f, p, r):
    from sovereign_solver_amd.engine.forms import extract
    x, _, y = extract(f, p, r["x"], r["y"])
    return x, y


@pytest.mark.parametrize("shape", ["transport", "random"])
def test_spmv_sell64_same_bits_as_one_lane_csr(gpu, shape, monkeypatch):
    """pdlp_spmv_sell64 over the engine's SELL-64 copy of K^T gives pdlp_spmv_csr's
    one-lane-per-row result bit for bit (each row summed in CSR order)."""
    import torch
    from sovereign_solver_amd.engine import _lib as L
    from sovereign_solver_amd.engine import sell as SL
    from sovereign_solver_amd.engine.device import ptr, stream_handle
    from sovereign_solver_amd.engine.forms import box_form
    from sovereign_solver_amd.engine.single import SingleEngine
    monkeypatch.setattr(SL, "COLUMN_PADDING_MAX", 100.0)
    p = transportation_lp(70, 150, seed=2) if shape == "transport" else random_sparse_lp(300, 700, 0.02, seed=4)
    eng = SingleEngine(box_form(p), "fp32")
    assert eng.t_sell_slices > 0
    rng = np.random.default_rng(0)
    x = torch.from_numpy(rng.standard_normal(eng.m).astype(np.float32)).to(gpu)
    a = torch.zeros(eng.n, dtype=torch.float32, device=gpu)
    b = torch.zeros(eng.n, dtype=torch.float32, device=gpu)
    sh = stream_handle(torch.cuda.current_stream())
    L.check(L.lib().pdlp_spmv_csr(L.PDLP_F32, ptr(eng.t_row_ptr), ptr(eng.t_col_idx), ptr(eng.t_val), ptr(x),
                                  ptr(a), eng.n, 1, sh), "spmv_csr")
    L.check(L.lib().pdlp_spmv_sell64(L.PDLP_F32, ptr(eng.t_sell_ptr), ptr(eng.t_sell_idx), ptr(eng.t_sell_val),
                                     ptr(x), ptr(b), eng.n, sh), "spmv_sell64")
    torch.cuda.synchronize()
    assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    want = (p.A_eq if p.n_ineq == 0 else sp.vstack([p.A_ub, p.A_eq])).T.tocsr() @ x.cpu().numpy().astype(np.float64)
    assert np.allclose(b.cpu().numpy(), want, rtol=1e-5, atol=1e-4)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_spmv_dev_over_slices_and_csr_rows(gpu, prec):
    """pdlp_spmv_dev: y = K x over an engine's SELL-64 rows and CSR rows in every
    one-launch layout (sell 2..7) and CSR-row width: within the rounding bound of
    the row's length against the FP64 product, and the same bits on a repeat."""
    import torch
    from sovereign_solver_amd.engine.device import ptr, stream_handle
    from sovereign_solver_amd.engine.forms import box_form
    from sovereign_solver_amd.engine.single import SingleEngine
    from sovereign_solver_amd.engine.tune import Tune
    p = transportation_lp(70, 150, seed=2)
    eng = SingleEngine(box_form(p), prec)
    assert eng.sell_slices > 0 and eng.n_csr_rows > 0
    dt = torch.float32 if prec == "fp32" else torch.float64
    rng = np.random.default_rng(1)
    x = torch.from_numpy(rng.standard_normal(eng.n)).to(gpu, dt)
    K = sp.csr_matrix((eng.val.double().cpu().numpy(), eng.col_idx.cpu().numpy(), eng.row_ptr.cpu().numpy()),
                      shape=(eng.m, eng.n))
    xd = x.double().cpu().numpy()
    want = K @ xd
    eps = 2.0**-24 if prec == "fp32" else 2.0**-53
    bound = (np.diff(K.indptr).max() + 2) * eps * (abs(K) @ abs(xd)) + 1e-300
    sh = stream_handle(torch.cuda.current_stream())
    for mode in (2, 3, 4, 5, 6, 7):
        for tc in (64, 256, 1024):
            t = Tune(tpr_primal=1, tpr_dual=tc, block=256, check_every=50, sell=mode).to_c()
            ys = []
            for _ in range(2):
                y = torch.full((eng.m,), float("nan"), dtype=dt, device=gpu)
                L.check(L.lib().pdlp_spmv_dev(C.byref(eng.dev), C.byref(t), ptr(x), ptr(y), sh), "spmv_dev")
                ys.append(y)
            torch.cuda.synchronize()
            a, b = ys[0].double().cpu().numpy(), ys[1].double().cpu().numpy()
            assert np.array_equal(a, b), (mode, tc)
            assert np.all(np.abs(a - want) <= bound), (mode, tc, np.abs(a - want).max())
    bad = Tune(tpr_primal=1, tpr_dual=64, block=256, check_every=50, sell=1).to_c()
    assert L.lib().pdlp_spmv_dev(C.byref(eng.dev), C.byref(bad), ptr(x), ptr(x), sh) == L.PDLP_E_TUNE


@pytest.mark.parametrize("shape", ["transport", "random"])
def test_sell_columns_from_device_widths(gpu, shape):
    """The SELL-64 copy's slice pointers from the device's slice widths
    (pdlp_sell64_widths over K^T's row pointers) equal the host's from K."""
    import torch
    from sovereign_solver_amd.engine import sell as SL
    from sovereign_solver_amd.engine.forms import box_form
    p = transportation_lp(70, 150, seed=2) if shape == "transport" else random_sparse_lp(300, 700, 0.02, seed=4)
    K = box_form(p).K
    KT = K.T.tocsr()
    s = torch.cuda.Stream()
    trp = torch.from_numpy(KT.indptr.astype(np.int32)).to(gpu)
    for pm in (10.0, 0.0):
        a = SL.build_columns(K, padding_max=pm)
        b = SL.build_columns_device(trp, K.nnz, s, padding_max=pm)
        assert (a is None) == (b is None)
        if a is not None:
            assert np.array_equal(a, b)


def test_uniform_values_read_none_same_bits(gpu):
    """A matrix whose entries are all one integer (a transportation LP: 1) with unpadded
    SELL copies runs the SELL steps without reading values (PdlpDev.k_const), and rows
    whose columns step by a constant without reading their indices past the first
    (PdlpDev.sell_stride / csr_stride): the same bits as reading them; a padded or
    non-uniform matrix keeps k_const = 0."""
    from sovereign_solver_amd.engine.forms import box_form
    from sovereign_solver_amd.engine.single import SingleEngine
    from sovereign_solver_amd.engine.tune import Tune
    p = transportation_lp(60, 200, seed=2)
    t = Tune(tpr_primal=1, tpr_dual=256, block=256, check_every=50, sell=6, sell_primal=1)
    eng = SingleEngine(box_form(p), "fp32", tune=t)
    assert eng.k_const == 1 and eng.sell_slices > 0 and eng.t_sell_slices > 0
    a = eng.solve(1e-4, 50_000)
    assert eng.sell_stride is not None and int(eng.sell_stride.min()) == 200     # demand rows step by 200
    assert eng.csr_stride is not None and int(eng.csr_stride.max()) >= 1
    eng.dev.k_const = 0
    eng.drop_graphs()
    b = eng.solve(1e-4, 50_000)
    eng.dev.sell_stride = eng.dev.csr_stride = None          # every index read as well
    eng.drop_graphs()
    c = eng.solve(1e-4, 50_000)
    assert a["status"] == b["status"] == c["status"] == "optimal"
    assert a["iterations"] == b["iterations"] == c["iterations"]
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["y"], b["y"])
    assert np.array_equal(a["x"], c["x"]) and np.array_equal(a["y"], c["y"])
    assert SingleEngine(box_form(random_sparse_lp(300, 700, 0.02, seed=4)), "fp32").k_const == 0


def test_device_row_strides_match_host(gpu):
    """pdlp_row_strides (the CSR rows' PdlpDev.csr_stride, found on the device) equals
    engine.sell.row_strides, which a direct check of each row pins, on rows with no, one,
    evenly and unevenly spaced entries, including rows longer than one wavefront."""
    from sovereign_solver_amd.engine import _lib as L
    from sovereign_solver_amd.engine.forms import box_form
    from sovereign_solver_amd.engine.sell import row_strides
    lib = L.lib()
    rows = []
    for j in range(200):                                     # lengths 0 .. 199, steps 0 .. 6
        n, st = j % 150, 1 + j % 7
        rows.append(list(range(0, n * st, st)))
    rows[5] = [0, 3, 6, 10]                                  # one uneven step at the end
    rows[150] = list(range(0, 300, 2)) + [301]               # past the first wavefront
    rows[151] = [7]
    ind = [np.array(r, np.int32) for r in rows]
    ptrs = np.concatenate([[0], np.cumsum([len(r) for r in ind])]).astype(np.int32)
    A = sp.csr_matrix((np.ones(ptrs[-1]), np.concatenate(ind), ptrs), shape=(len(rows), 1200))
    for K in (A, box_form(transportation_lp(60, 200, seed=2)).K,
              box_form(random_sparse_lp(300, 700, 0.02, seed=4)).K):
        sel = np.arange(K.shape[0], dtype=np.int32)[::-1].copy()
        want = row_strides(K, sel)
        for r, s_ in zip(sel, want):
            idx = K.indices[K.indptr[r]: K.indptr[r + 1]]
            assert s_ == (0 if idx.size == 0 else 1 if idx.size == 1 else
                          (idx[1] - idx[0] if np.array_equal(idx, idx[0] + (idx[1] - idx[0]) * np.arange(idx.size))
                           else 0))
        d = [torch.from_numpy(a).to(gpu) for a in (sel, K.indptr.astype(np.int32), K.indices.astype(np.int32))]
        out = torch.full((sel.size,), -7, dtype=torch.int32, device=gpu)
        assert lib.pdlp_row_strides(int(sel.size), d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(),
                                    out.data_ptr(), None) == 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), want)
