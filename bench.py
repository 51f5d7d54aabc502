"""Benchmark: LPs/s to 1e-4 relative KKT on one MI355X — learned + batched vs
plain PDHG — with the roofline of the dominant kernel and the CPU baseline.

Default run (``python bench.py [--steps K --warmup W]``): the headline,
config 3 at its full size — economic dispatch, 200 generators x 24 periods
(m = 9,224 rows, n = 4,800 columns), a 3-layer width-512 MLP trained on
50,000 PDHG-labelled instances, 1,024 held-out instances solved in one
shared-matrix batch warm-started by the network (x and primal weight), every
solve finished and judged by PDHG at relative KKT <= 1e-4 (reference formula
on the reference-form lift, FP64).  A *step* is one such solve of all 1,024
LPs with their data resident in HBM: network inference, device-side start,
the batched solve and the results gathered in HBM.  ``value`` is LPs/s over
the K timed steps.  The same run measures, on the same held-out LPs:

* the host-to-host rate (problems on the host -> features, device packing,
  inference, solve, results and objectives on the host);
* the cold-started batched engine, and plain single-instance GPU PDHG (one
  LP after another, cold start) both with its LP resident and through the
  drop-in call ``solve_lp_pdhg`` (``speedup_vs_plain``);
* the objective error of 64 LPs against HiGHS dual simplex (FP64);
* ``cpu_baseline``: the FP64 CPU restatement of the reference's PDHG on a
  pool of the host's cores (and on one core), on LPs of the same set.

``config.subrecords`` carry the other configs, each with its clock: C1 the
transportation LP (1,000 x 2,000), C2 the 4,096 uneven LPs (two persistent
launches), C4 256 set-cover node relaxations, and ``beyond_mall`` the
standalone SpMV and fused steps on a transportation LP whose working set
exceeds the 256 MB Infinity Cache.  ``--workload`` runs one part alone.
"""

from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
METRIC = "LPs/sec to 1e-4 rel KKT: learned+batched vs plain PDHG, 1 GPU; SpMV HBM GB/s"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(key: str):
    """HBM bytes per launch of a kernel on this workload, as measured by the
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs recorded in profiles/traffic.json
    (None when that workload was not profiled)."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            rec = json.load(f).get(key)
    except (OSError, ValueError):
        return None
    return None if rec is None else float(rec["bytes"])


# ------------------------------------------------------------------ references
def highs_objective(p, method: str = "highs-ds", time_limit: float = 0.0):
    """FP64 optimum of p by HiGHS (scipy; dual simplex, or interior point with
    crossover), or None when HiGHS does not finish (within ``time_limit`` s if > 0)."""
    from scipy.optimize import linprog
    opts = {"time_limit": float(time_limit)} if time_limit > 0 else {}
    r = linprog(p.c, A_ub=p.A_ub if p.n_ineq else None, b_ub=p.b_ub if p.n_ineq else None,
                A_eq=p.A_eq if p.n_eq else None, b_eq=p.b_eq if p.n_eq else None,
                bounds=list(zip([None if np.isinf(v) else v for v in p.lb],
                                [None if np.isinf(v) else v for v in p.ub])), method=method, options=opts)
    if r.status != 0:
        return None
    return -float(r.fun) if p.sense == "max" else float(r.fun)


def objective_errors(probs, objs, k: int, method: str = "highs-ds", time_limit: float = 0.0,
                     total_cap_s: float = 120.0):
    """|obj - obj*| / (1 + |obj*|) on the first k LPs, obj* by HiGHS (FP64); LPs that
    HiGHS does not finish within its per-LP limit, or after the total cap, are counted
    as unchecked."""
    t0 = time.perf_counter()
    errs, unchecked = [], 0
    for p, o in list(zip(probs, objs))[:k]:
        if time.perf_counter() - t0 > total_cap_s:
            unchecked += 1
            continue
        ref = highs_objective(p, method, time_limit)
        if ref is None:
            unchecked += 1
        else:
            errs.append(abs(o - ref) / (1.0 + abs(ref)))
    e = np.array(errs) if errs else np.array([np.nan])
    return {"lps": len(errs), "unchecked": unchecked, "max": float(e.max()), "median": float(np.median(e)),
            "reference": f"HiGHS {method} (scipy.optimize.linprog, FP64)"
                         + (f", {time_limit:g} s per LP" if time_limit > 0 else ""),
            "s": time.perf_counter() - t0}


# ---------------------------------------------------------------- CPU baseline
_CPU_CACHE = {}


def _cpu_init():
    os.environ["OMP_NUM_THREADS"] = "1"
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    try:
        from threadpoolctl import threadpool_limits
        threadpool_limits(1)
    except Exception:                                   # pragma: no cover
        pass


def _cpu_problem(spec):
    """The LP of a workload spec, rebuilt from its seeds (no data crosses processes)."""
    from sovereign_solver_amd.workloads import DispatchFamily, SetCoverNodes, random_sparse_lp, uneven_sizes
    kind, k = spec
    if kind == "dispatch":
        fam = _CPU_CACHE.get(kind) or _CPU_CACHE.setdefault(kind, DispatchFamily())
        return fam.problem(k)
    if kind.startswith("uneven/"):                       # "uneven/<batch size>"
        ms = _CPU_CACHE.get(kind)
        if ms is None:
            ms = _CPU_CACHE.setdefault(kind, uneven_sizes(int(kind.split("/")[1]), 50, 2000, 0))
        m = int(ms[k])
        return random_sparse_lp(m, int(round(2.5 * m)), 0.01, seed=k, family=2)
    if kind == "setcover":
        sc = _CPU_CACHE.get(kind) or _CPU_CACHE.setdefault(kind, SetCoverNodes())
        return sc.node(k)
    raise ValueError(kind)


def _cpu_noop(_):
    return os.getpid()


def _cpu_solve(job):
    from oracle import pdhg_ref
    spec, tol = job
    p = _cpu_problem(spec)
    t0 = time.perf_counter()
    r = pdhg_ref.solve(p, max_iters=100_000, tol=tol)
    return time.perf_counter() - t0, r.status


def cpu_workers() -> int:
    """The host cores this process may use, at most 16 (one GPU's share of a box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:                              # pragma: no cover
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _pool_rate(jobs, workers: int, cap_s: float):
    """LPs finished per second by ``workers`` single-threaded processes in cap_s seconds
    of wall time (LPs still running at the deadline count for nothing)."""
    ctx = mp.get_context("spawn")
    pool = ctx.Pool(workers, initializer=_cpu_init)
    try:
        pool.map(_cpu_noop, range(workers))              # interpreters started before the clock
        t0 = time.perf_counter()
        deadline = t0 + cap_s
        done = 0
        it = pool.imap_unordered(_cpu_solve, jobs)
        while True:
            left = deadline - time.perf_counter()
            if left <= 0:
                break
            try:
                it.next(timeout=left)
            except (StopIteration, mp.TimeoutError):
                break
            done += 1
        wall = min(time.perf_counter(), deadline) - t0
    finally:
        pool.terminate()
        pool.join()
    return done, wall


def cpu_baseline(kind: str, ids, tol: float, cap_s: float, what: str):
    """The FP64 CPU restatement of the reference's PDHG (``oracle/pdhg_ref.py``) on
    LPs of the workload, in a random order: a pool of single-threaded processes over
    the host cores for cap_s seconds of wall time, and one process for cap_s / 2."""
    _cpu_init()
    workers = cpu_workers()
    jobs = [((kind, int(k)), tol) for k in np.random.default_rng(11).permutation(np.asarray(ids))]
    done, wall = _pool_rate(jobs, workers, cap_s)
    pool_rate = done / wall
    one, wall1 = _pool_rate(jobs, 1, cap_s / 2)
    one_rate = one / wall1
    last, t0 = wall, 0.0
    try:
        model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
    except (OSError, StopIteration):                    # pragma: no cover
        model = "unknown"
    log(f"cpu baseline ({what}): {pool_rate:.3f} LPs/s on {workers} processes ({done} LPs), "
        f"{one_rate:.3f} LPs/s on 1 core ({one} LPs)")
    return {"value": pool_rate, "unit": "LPs/s", "cores": workers, "kind": "port",
            "sample": (f"oracle FP64 reference-form PDHG (restatement of solver/lp/pdhg.py, SciPy/NumPy) on {what}, "
                       f"to {tol:g}, in a seeded random order: a pool of {workers} single-threaded processes "
                       f"finished {done} LPs in {last - t0:.1f}s of wall time (LPs unfinished at the deadline "
                       f"count for nothing); one process alone {one_rate:.3f} LPs/s ({one} LPs); "
                       f"OMP/BLAS threads 1; host {model}, cpu_count {os.cpu_count()}"),
            "one_core_lps_per_s": one_rate}


# --------------------------------------------------------------------- timing
def primal_bytes(form, elem: int, sell_entries: int = 0, values: bool = True) -> float:
    """Algorithmic bytes of one fused primal-step launch.  K^T as CSR (values, row
    indices, column pointers) or, with ``sell_entries`` > 0, as the SELL-64 copy the
    step then reads (values and row indices of every slot, padding included, and a
    pointer per slice of 64 columns, no column pointers; no values when ``values`` is
    false: every entry holds one constant, PdlpDev.k_const)."""
    n, m, nnz = form.n, form.m, form.nnz
    if sell_entries:
        b = sell_entries * ((elem if values else 0) + 4) + (-(-n // 64) + 1) * 4
    else:
        b = nnz * (elem + 4) + (n + 1) * 4    # K^T CSR: values, column indices, row pointers
    b += m * elem                             # y, gathered
    b += n * elem                             # c
    if form.lo is not None:
        b += 2 * n * elem                     # lo, hi
    b += 2 * n * elem                         # x read + write
    b += n * elem                             # xbar write
    b += 2 * n * 4                            # FP32 partial sum read + write
    return float(b)


def pkern(eng) -> str:
    """Name of the primal-step kernel the engine's tuning launches."""
    return {0: "primal_kernel", 1: "primal_sell_kernel", 2: "primal_sell2_kernel"}[eng.tune.sell_primal]


def dual_bytes(form, elem: int, values: bool = True, index_loads=None) -> float:
    """Algorithmic bytes of one fused dual-step launch (K's values only when ``values``;
    ``index_loads`` column indices, default one per entry: a row whose columns step by
    a constant reads only its first)."""
    n, m, nnz = form.n, form.m, form.nnz
    il = nnz if index_loads is None else index_loads
    b = nnz * (elem if values else 0) + il * 4 + (m + 1) * 4 + n * elem   # K CSR + xbar gathered once
    b += m * elem + 2 * m * elem + 2 * m * 4        # b, y r/w, FP32 partial sum r/w
    if form.row_kind is not None:
        b += m
    return float(b)


def spmv_rates(eng, reps: int = 50, kx_tpr: int = 64, traffic_key=None):
    """Standalone FP32 SpMV (``pdlp_spmv_csr``) on the engine's K and K^T: a HIP
    graph of ``reps`` back-to-back launches timed with events on its stream.
    Algorithmic bytes per product: 8 nnz + 4 (rows + 1) + 4 cols + 4 rows (FP32
    values, int32 indices, x gathered once, y written once)."""
    import ctypes as C
    import torch
    from sovereign_solver_amd.engine import _lib as L
    from sovereign_solver_amd.engine.device import ptr, stream_handle
    lib = eng.lib
    out = {}
    s = torch.cuda.Stream(device=eng.device)
    cases = [("KT_y", eng.t_row_ptr, eng.t_col_idx, eng.t_val, eng.y, eng.n, eng.m, 1),
             ("K_x", eng.row_ptr, eng.col_idx, eng.val, eng.x, eng.m, eng.n, kx_tpr)]
    if getattr(eng, "t_sell_slices", 0):       # K^T y over the SELL-64 copy of K^T (pdlp_spmv_sell64)
        cases.insert(1, ("KT_y_sell", eng.t_sell_ptr, eng.t_sell_idx, eng.t_sell_val, eng.y, eng.n, eng.m, 0))
    tune_c = None
    if getattr(eng, "sell_slices", 0) and 2 <= eng.tune.sell <= 7:
        # K x over the engine's SELL-64 rows and CSR rows in its tuned one-launch layout (pdlp_spmv_dev)
        cases.append(("K_x_sell", eng.row_ptr, eng.col_idx, eng.val, eng.x, eng.m, eng.n, -1))
        tune_c = eng.tune.to_c()
    for name, rp, ci, val, x, rows, cols, tpr in cases:
        with torch.cuda.stream(s):
            y = torch.empty(rows, dtype=torch.float32, device=eng.device)

        def launch():
            if tpr == -1:
                L.check(lib.pdlp_spmv_dev(C.byref(eng.dev), C.byref(tune_c), ptr(x), ptr(y),
                                          stream_handle(torch.cuda.current_stream())), "pdlp_spmv_dev")
            elif tpr == 0:
                L.check(lib.pdlp_spmv_sell64(L.PDLP_F32, ptr(rp), ptr(ci), ptr(val), ptr(x), ptr(y), rows,
                                             stream_handle(torch.cuda.current_stream())), "pdlp_spmv_sell64")
            else:
                L.check(lib.pdlp_spmv_csr(L.PDLP_F32, ptr(rp), ptr(ci), ptr(val), ptr(x), ptr(y), rows, tpr,
                                          stream_handle(torch.cuda.current_stream())), "pdlp_spmv_csr")
        with torch.cuda.stream(s):
            launch()
        s.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            for _ in range(reps):
                launch()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(s):
            g.replay()                                 # warm
            e0.record(s)
            g.replay()
            e1.record(s)
        s.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        nnz = int(val.numel())                   # SELL: slots, padding included
        ptrs = (int(rp.numel()) if tpr == 0 else rows + 1)
        nbytes = 8 * nnz + 4 * ptrs + 4 * cols + 4 * rows
        if tpr == -1:                            # the dual step's layout: index runs, constant values
            nbytes -= 4 * (nnz - eng.dual_index_loads)
            if eng.k_const:
                nbytes -= 4 * nnz
        gbs = nbytes / (us * 1e-6) / 1e9
        out[name] = {"rows": rows, "nnz": nnz, "lanes_per_row": tpr if tpr >= 0 else "sell_wg", "bytes": nbytes,
                     "us": us,
                     "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS,
                     "traffic": measured_traffic(f"{traffic_key}/spmv_{name}") if traffic_key else None}
        del g
    return out


def family_passes(eng, tol, traffic_key=None, reps: int = 20):
    """Graph-timed column and row passes of a family engine over all its LPs at the
    cold-start state (every LP running): algorithmic bytes / mean time."""
    fb = eng.fb
    eng.start(tol)
    us = {w: eng.time_step(w, reps=reps) for w in ("primal", "dual")}
    b = {"primal": fb.column_pass_bytes(), "dual": fb.row_pass_bytes()}
    gbs = {w: b[w] / (us[w] * 1e-6) / 1e9 for w in us}
    dom = max(us, key=us.get)
    names = {"primal": "family column pass (primal_kernel: long and short columns in one launch)",
             "dual": "family row pass (dual_kernel: long and short rows in one launch)"}
    log(f"family passes: column {us['primal']:.1f} us ({b['primal'] / 1e6:.1f} MB, {gbs['primal']:.0f} GB/s), "
        f"row {us['dual']:.1f} us ({b['dual'] / 1e6:.1f} MB, {gbs['dual']:.0f} GB/s)")
    other = "dual" if dom == "primal" else "primal"
    return {"bound": "hbm", "kernel": names[dom], "achieved": gbs[dom], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": gbs[dom] / HBM_PEAK_GBS,
            "traffic": measured_traffic(f"{traffic_key}/{'column' if dom == 'primal' else 'row'}_pass")
            if traffic_key else None,
            "bytes_per_launch": b[dom], "us_per_launch": us[dom],
            "launch_shape": {"items_per_wave": eng.args.items_per_wave, "xcd_affinity": eng.args.xcd_affinity},
            "other_pass": {"kernel": names[other], "achieved": gbs[other], "bytes_per_launch": b[other],
                           "us_per_launch": us[other],
                           "traffic": measured_traffic(f"{traffic_key}/{'row' if other == 'dual' else 'column'}"
                                                       f"_pass") if traffic_key else None}}


def node_forward_rate(model, V_dev, reps: int = 10):
    """MFMA rate of the node network on the instances V_dev: every stage's gather, three
    GEMMs (2 rows (d_in d_h + d_h d_h + d_h d_out) FLOP each) and the assembly, timed
    with HIP events on the network's stream over ``reps`` back-to-back passes."""
    from sovereign_solver_amd.learned import nodes as N
    st = model.mlp.stream
    for _ in range(2):
        model.starts(V_dev)
    e0, e1 = torch_events()
    e0.record(st)
    for _ in range(reps):
        model.starts(V_dev)
    e1.record(st)
    st.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / reps
    rows = V_dev.shape[0] * model.spec.nodes
    flop = N.forward_flop(model, rows)
    tflops = flop / (us * 1e-6) / 1e12
    return {"instances": V_dev.shape[0], "rows": rows, "gflop": flop / 1e9, "us": us, "tflops": tflops,
            "peak_tflops": BF16_PEAK_TFLOPS, "frac": tflops / BF16_PEAK_TFLOPS, "bound": "mfma (dense BF16)"}


def torch_events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def plain_rates(probs, tol, max_iters):
    """Plain single-instance GPU PDHG (cold start), one LP after another: LPs/s with
    the LP resident in HBM (engine built before the clock, solve only) and through the
    drop-in call ``solve_lp_pdhg`` (host problem to host result)."""
    import torch
    from sovereign_solver_amd import solve_lp_pdhg
    from sovereign_solver_amd.engine.forms import box_form
    from sovereign_solver_amd.engine.single import SingleEngine
    from sovereign_solver_amd.engine._lib import PDLP_SCHEME_ADAPTIVE
    t_res, ok_res, t_call, ok_call = 0.0, 0, 0.0, 0
    its = []
    for p in probs:
        # the drop-in's own default for the box form: adaptive restarts and primal weight
        se = SingleEngine(box_form(p), "fp32", scheme=PDLP_SCHEME_ADAPTIVE)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = se.solve(tol, max_iters)
        t_res += time.perf_counter() - t0
        ok_res += r["status"] == "optimal"
        its.append(r["iterations"])
        se.drop_graphs()
        del se
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rc = solve_lp_pdhg(p, max_iters=max_iters, tol=tol, precision="fp32")
        t_call += time.perf_counter() - t1
        ok_call += rc.status == "optimal"
    return {"lps": len(probs), "resident_lps_per_s": ok_res / t_res if t_res else 0.0,
            "drop_in_call_lps_per_s": ok_call / t_call if t_call else 0.0,
            "optimal": ok_res, "iters_mean": float(np.mean(its)), "iters_max": int(np.max(its))}


def calib_copy(args):
    """Three 1 GiB device-to-device copies: known bytes for the HBM counter scale."""
    if not args.calib_copy:
        return
    import torch
    a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log("calibration copies: 3 x 1 GiB read + 1 GiB written")


# -------------------------------------------------------------- C3: headline
def run_headline(args):
    """Config 3: learned warm start + shared-matrix batched PDHG on the economic
    dispatch family.  Returns (JSON line fields, roofline, cpu baseline)."""
    import torch
    from sovereign_solver_amd.engine.family import FamilyBatch, FamilyEngine
    from sovereign_solver_amd.learned import nodes as N
    from sovereign_solver_amd.learned import warmstart as W
    from sovereign_solver_amd.workloads import DispatchFamily

    tol, mi = args.tol, args.max_iters
    fam = DispatchFamily()
    n, m = fam.n, fam.m
    log(f"dispatch family 200x24: m={m} n={n} nnz={fam.A.nnz}; training on {args.train}, {args.eval} held out")
    train_ids = np.arange(args.train)
    eval_ids = 10**6 + np.arange(args.eval)
    t0 = time.perf_counter()
    lab = W.label(fam, train_ids, tol=tol, max_iters=mi)
    t_label = time.perf_counter() - t0
    keep = lab.status == "optimal"
    log(f"labelled {len(train_ids)} in {t_label:.1f}s (engine {lab.seconds:.1f}s), optimal {int(keep.sum())}, "
        f"iterations mean {lab.iterations.mean():.0f}")
    t1 = time.perf_counter()
    # the weight is calibrated on 256 training instances (one scalar for the family)
    model = N.train_node_model(fam, lab, stages=args.stages, steps=args.node_steps, rows=args.node_rows,
                               calib_ids=train_ids[-256:], tol=tol, max_iters=mi, log=log)
    t_train = time.perf_counter() - t1
    stages = N.stage_list(model)
    train_flop = sum(s.train_flop for s in stages)
    train_gpu_s = sum(s.train_gpu_seconds for s in stages)
    train_rate = {"tflops": train_flop / train_gpu_s / 1e12, "frac": train_flop / train_gpu_s / 1e12 / BF16_PEAK_TFLOPS,
                  "gpu_s": train_gpu_s, "rows_per_step": args.node_rows, "steps_per_stage": args.node_steps}
    log(f"trained {len(stages)} node-model stages in {t_train:.1f}s (label excluded; device step loop "
        f"{train_gpu_s:.1f}s at {train_rate['tflops']:.0f} TFLOP/s, {100 * train_rate['frac']:.1f}% of dense BF16), "
        f"primal weight {'default' if model.log_weight is None else f'{np.exp(model.log_weight):.2f}'}")

    probs = [fam.problem(int(k)) for k in eval_ids]
    t2 = time.perf_counter()
    eng = FamilyEngine(FamilyBatch(probs))
    shape = eng.autotune(tol)
    t_setup = time.perf_counter() - t2
    log(f"engine for {len(probs)} LPs built in {t_setup:.2f}s (once per family and batch size), shape {shape}")
    with torch.cuda.stream(model.mlp.stream):
        V_dev = torch.from_numpy(fam.node_inputs_from(probs)).to(model.device)
    model.mlp.stream.synchronize()
    net_rate = node_forward_rate(model, V_dev)
    log(f"node network (both stages + gathers + assembly) on {len(probs)} LPs ({net_rate['rows']} rows per stage): "
        f"{net_rate['gflop']:.0f} GFLOP in {net_rate['us']:.0f} us ({net_rate['tflops']:.0f} TFLOP/s, "
        f"{100 * net_rate['frac']:.1f}% of dense BF16 peak)")
    w_col = n + m if model.log_weight is not None else -1
    out_host = torch.empty(eng.t["out_final"].shape, dtype=torch.uint8).pin_memory()
    from sovereign_solver_amd.engine.batch import OUT_DTYPE

    def records():
        return np.frombuffer(out_host.numpy().tobytes(), dtype=OUT_DTYPE).copy()

    def learned_resident():
        out, ev = model.starts(V_dev)
        eng.start_from_device(tol, out, x_off=0, y_off=n, w_col=w_col, ready=ev)
        eng.run(mi)
        eng.collect()
        with torch.cuda.stream(eng.stream):
            out_host.copy_(eng.t["out_final"], non_blocking=True)

    def cold_resident():
        eng.start(tol)
        eng.run(mi)
        eng.collect()
        with torch.cuda.stream(eng.stream):
            out_host.copy_(eng.t["out_final"], non_blocking=True)

    def timed(fn, steps):
        torch.cuda.synchronize()
        ok, secs, last = 0, 0.0, None
        for _ in range(steps):
            t = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            secs += time.perf_counter() - t
            last = records()
            ok += int((last["status"] == 1).sum())
        return ok, secs, last

    # restart-check cadence by time to tolerance of the learned step itself (autotuner's search)
    from sovereign_solver_amd.engine import tune as T
    cadence = {"check_every": eng.check_every, "trials": {}}
    if not args.no_autotune:
        def measure(cfg):
            eng.set_check_every(cfg["check_every"])
            learned_resident()
            o, s, r = timed(learned_resident, 1)
            return {"seconds": s, "ok": o == len(probs), "iters_mean": float(r["iterations"].mean())}
        best, trials = T.coordinate_search({"check_every": (25, 50, 100)}, {"check_every": 50}, measure)
        eng.set_check_every(best["check_every"])
        cadence = {"check_every": best["check_every"], "trials": trials}
        log(f"check cadence: {best['check_every']} ({ {k: round(v['seconds'] * 1e3, 1) for k, v in trials.items()} } ms)")
    for _ in range(args.warmup):
        learned_resident()
    torch.cuda.synchronize()
    ok, secs, rec = timed(learned_resident, args.steps)
    value = ok / secs
    log(f"learned + batched, resident, batch by batch: {args.steps} steps x {len(probs)} LPs, {ok} optimal in {secs:.3f}s -> "
        f"{value:.0f} LPs/s (iterations mean {rec['iterations'].mean():.0f}, max {rec['iterations'].max()}, "
        f"device rel KKT max {rec['rel_kkt'].max():.2e})")
    learned = {"lps_per_s": value, "ms_per_step": secs * 1e3 / args.steps, "optimal_per_step": ok / args.steps,
               "iters_mean": float(rec["iterations"].mean()), "iters_max": int(rec["iterations"].max()),
               "device_rel_kkt_max": float(rec["rel_kkt"].max())}
    cold_resident()
    ok_c, secs_c, rec_c = timed(cold_resident, max(2, min(args.steps, 5)))
    cold = {"lps_per_s": ok_c / secs_c, "iters_mean": float(rec_c["iterations"].mean()),
            "iters_max": int(rec_c["iterations"].max())}
    log(f"cold batched, resident: {cold['lps_per_s']:.0f} LPs/s (iterations mean {cold['iters_mean']:.0f}, "
        f"max {cold['iters_max']})")
    # iterations at the reference's check cadence (50, config.py:38), where an LP can stop at any
    # multiple of 50: the learned start's gain in iterations without the tuned cadence's rounding
    ce_tuned = eng.check_every
    eng.set_check_every(50)
    at50 = {}
    for name, fn in (("learned", learned_resident), ("cold", cold_resident)):
        _, _, r50 = timed(fn, 1)
        at50[name] = {"iters_mean": float(r50["iterations"].mean()), "iters_max": int(r50["iterations"].max())}
    eng.set_check_every(ce_tuned)
    cold["iterations_at_cadence_50"] = at50

    # the same steps streamed through the engine (FamilyEngine.run_stream): a step's LPs enter
    # tiles as earlier steps' LPs finish, so the device runs full tiles instead of waiting on
    # each step's slowest LPs; every LP's result is the batch solve's bit for bit
    from sovereign_solver_amd.engine.family import StreamBatch
    src_dev = {k: eng.staged(k) for k in ("c", "b", "lo", "hi") if eng.staged(k) is not None}
    mi_s = -(-mi // eng.check_every) * eng.check_every

    def stream_run(steps, learned_start):
        bs = []
        for _ in range(steps):
            if learned_start:
                out, ev = model.starts(V_dev)
                bs.append(StreamBatch(src_dev, start=out, ready=ev, x_off=0, y_off=n, w_col=w_col))
            else:
                bs.append(StreamBatch(src_dev))
        O, _, _ = eng.run_stream(tol, mi_s, bs)
        return np.frombuffer(O.cpu().numpy().tobytes(), dtype=OUT_DTYPE)

    streamed = {}
    for name, ls in (("learned", True), ("cold", False)):
        stream_run(2, ls)                                   # graphs, buffers
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = stream_run(args.steps, ls)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        ok_s = int((r["status"] == 1).sum())
        streamed[name] = {"lps_per_s": ok_s / dt, "ms_per_step": dt * 1e3 / args.steps, "steps": args.steps,
                          "optimal": ok_s, "iters_mean": float(r["iterations"].mean()),
                          "iters_max": int(r["iterations"].max()), "device_rel_kkt_max": float(r["rel_kkt"].max())}
    log(f"streamed, resident: learned {streamed['learned']['lps_per_s']:.0f} LPs/s "
        f"({streamed['learned']['ms_per_step']:.1f} ms per step, {streamed['learned']['optimal']} optimal of "
        f"{args.steps * len(probs)}), cold {streamed['cold']['lps_per_s']:.0f} LPs/s")
    value = streamed["learned"]["lps_per_s"]
    log(f"iterations at cadence 50: learned mean {at50['learned']['iters_mean']:.0f} (max "
        f"{at50['learned']['iters_max']}), cold mean {at50['cold']['iters_mean']:.0f} (max {at50['cold']['iters_max']}) "
        f"-> {at50['cold']['iters_mean'] / at50['learned']['iters_mean']:.2f}x")

    # host to host: problems on the host -> staging and upload, network inputs (on the device,
    # from the staged vectors), inference, solve -> results and objectives; one batch at a time,
    # then a serving loop over consecutive batches whose host work overlaps the device work
    from sovereign_solver_amd.lp.serve import FamilyServer
    kw = {"items_per_wave": shape["items_per_wave"], "xcd_affinity": bool(shape["xcd_affinity"]),
          "check_every": eng.check_every}
    servers = {"learned": FamilyServer(probs, tol, mi, model=model, features=fam.node_inputs_staged, engine_kw=kw),
               "cold": FamilyServer(probs, tol, mi, engine_kw=kw)}
    h2h = {}
    for name, srv in servers.items():
        srv.solve(probs)                              # graphs captured, buffers warm
        torch.cuda.synchronize()
        t = time.perf_counter()
        reps = 3
        for _ in range(reps):
            o, X, Y, obj = srv.solve(probs)
        dt = (time.perf_counter() - t) / reps
        h2h[name] = {"lps_per_s": int((o["status"] == 1).sum()) / dt, "ms_per_batch": dt * 1e3}
        if name == "learned":
            learned_obj = obj.copy()
    log(f"host to host: learned {h2h['learned']['lps_per_s']:.0f} LPs/s ({h2h['learned']['ms_per_batch']:.1f} ms "
        f"per batch), cold {h2h['cold']['lps_per_s']:.0f} LPs/s ({h2h['cold']['ms_per_batch']:.1f} ms)")
    # one more learned batch with a synchronisation after each stage: where the time goes
    srv = servers["learned"]
    e0 = srv.engines[0]
    stages = {}
    torch.cuda.synchronize()
    t = time.perf_counter()
    for name, fn in (("load", lambda: e0.load(probs)),
                     ("inputs", lambda: fam.node_inputs_staged(probs, e0)),
                     ("inference", lambda: model.starts(V)),
                     ("start_solve", lambda: (e0.start_from_device(tol, S0[0], x_off=0, y_off=n,
                                                                  w_col=n + m if model.log_weight is not None
                                                                  else -1, ready=S0[1]), e0.run(mi))),
                     ("results", lambda: e0.results()),
                     ("objectives", lambda: e0.objectives(X))):
        r = fn()
        if name == "inputs":
            V = r
        if name == "inference":
            S0 = r
        if name == "results":
            X = r[1]
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        stages[name] = (t1 - t) * 1e3
        t = t1
    h2h["learned"]["stages_ms"] = stages
    log("host to host, learned, by stage (ms): " + ", ".join(f"{k} {v:.1f}" for k, v in stages.items()))
    nb = args.serve_batches
    batches = [[fam.problem(int(10**6 + len(probs) * (j + 1) + k)) for k in range(len(probs))] for j in range(nb)]
    for name, srv in servers.items():
        torch.cuda.synchronize()
        t = time.perf_counter()
        ok = sum(int((o["status"] == 1).sum()) for o, X, Y, obj in srv.solve_stream(batches))
        dt = time.perf_counter() - t
        h2h[name]["pipelined"] = {"lps_per_s": ok / dt, "ms_per_batch": dt * 1e3 / nb, "batches": nb,
                                  "optimal": ok}
    # the same batches through one engine as one stream, host to host (FamilyServer.solve_streamed)
    for name, srv in servers.items():
        srv.solve_streamed(batches[:1])                 # graphs, buffers
        torch.cuda.synchronize()
        t = time.perf_counter()
        ok = sum(int((o["status"] == 1).sum()) for o, X, Y, obj in srv.solve_streamed(batches))
        dt = time.perf_counter() - t
        h2h[name]["streamed"] = {"lps_per_s": ok / dt, "ms_per_batch": dt * 1e3 / nb, "batches": nb,
                                 "optimal": ok}
    log(f"host to host, streamed through one engine over {nb} batches: learned "
        f"{h2h['learned']['streamed']['lps_per_s']:.0f} LPs/s, cold {h2h['cold']['streamed']['lps_per_s']:.0f} LPs/s")
    del servers, srv, e0
    log(f"host to host, serving loop over {nb} batches: learned {h2h['learned']['pipelined']['lps_per_s']:.0f} "
        f"LPs/s ({h2h['learned']['pipelined']['ms_per_batch']:.1f} ms per batch), cold "
        f"{h2h['cold']['pipelined']['lps_per_s']:.0f} LPs/s ({h2h['cold']['pipelined']['ms_per_batch']:.1f} ms)")
    roof = family_passes(eng, tol, f"dispatch_{len(probs)}")
    del eng
    plain = plain_rates(probs[: args.plain_sample], tol, mi)
    log(f"plain single-instance GPU PDHG on {plain['lps']} of the held-out LPs: resident "
        f"{plain['resident_lps_per_s']:.2f} LPs/s, drop-in call {plain['drop_in_call_lps_per_s']:.2f} LPs/s")
    oe = objective_errors(probs, learned_obj, args.obj_sample)
    log(f"objective vs HiGHS on {oe['lps']} LPs: max {oe['max']:.2e}, median {oe['median']:.2e} ({oe['s']:.1f}s)")
    cpu = None if args.no_cpu_baseline else cpu_baseline("dispatch", eval_ids, tol, args.cpu_cap_s,
                                                         "held-out dispatch LPs")
    fields = {
        "value": value, "ms_per_step": streamed["learned"]["ms_per_step"],
        "dtype": "fp32 iterates and SpMV, fp64 KKT and objectives; network bf16 x bf16 -> fp32",
        "data": "synthetic: seeded economic-dispatch family (random-init network trained on PDHG's own solutions)",
        "config": {
            "workload": (f"C3 learned warm start + shared-matrix batched restarted PDHG, economic dispatch 200x24 "
                         f"(m={m}, n={n}), {len(probs)} held-out LPs per step, tol {tol:g}; step = network "
                         f"inference + device-side start + batched solve, inputs and results in HBM; the "
                         f"steps streamed through one engine (a step's LPs enter tiles as earlier ones finish)"),
            "lps_per_step": len(probs), "train_instances": args.train,
            "network": (f"node model: {len(stages)} stages of a 3-layer width-512 MLP shared by all 4,800 "
                        f"variables (second stage reads the first's outputs at t-2..t+2)"),
            "train_label_s": t_label, "train_s": t_train, "train_mfma": train_rate,
            "primal_weight": None if model.log_weight is None else float(np.exp(model.log_weight)),
            "engine_setup_s": t_setup, "launch_shape": shape, "check_cadence": cadence,
            "network_forward_mfma": net_rate,
            "learned_streamed_resident": streamed["learned"], "cold_streamed_resident": streamed["cold"],
            "learned_batched_resident": learned, "cold_batched_resident": cold,
            "host_to_host": h2h, "plain_single_gpu": plain,
            "speedup_vs_plain": value / plain["resident_lps_per_s"] if plain["resident_lps_per_s"] else None,
            "speedup_vs_plain_host_to_host": (h2h["learned"]["lps_per_s"] / plain["drop_in_call_lps_per_s"]
                                              if plain["drop_in_call_lps_per_s"] else None),
            "learned_vs_cold_batched": value / streamed["cold"]["lps_per_s"],
            "objective_rel_err_vs_fp64": oe,
        },
    }
    return fields, roof, cpu


# ---------------------------------------------------------- C1: transport LP
def run_transport(args, sources=1000, sinks=2000, cpu=True):
    """Config 1: one transportation LP, plain restarted PDHG (cold start)."""
    import torch
    from sovereign_solver_amd.engine.forms import box_form, extract, reference_form
    from sovereign_solver_amd.engine.single import SingleEngine
    from sovereign_solver_amd.engine import tune as T
    from sovereign_solver_amd.workloads import transportation_lp

    tol, mi = args.tol, args.max_iters
    p = transportation_lp(sources, sinks, seed=0)
    form = box_form(p)
    eng = SingleEngine(form, "fp32")
    tune = T.autotune(eng) if not args.no_autotune else eng.tune
    eng.set_tune(tune)
    cadence = None
    if not args.no_autotune:                 # check cadence by time to tolerance (one solve per candidate)
        tune = T.tune_cadence(eng, tol, mi, use_cache=False)
        cadence = T._load_cache().get(T._bucket(eng.m, eng.n, eng.nnz, eng.precision), {}).get("cadence")
        if cadence:
            log(f"check cadence {tune.check_every}: " + ", ".join(
                f"{k}: {v['seconds']:.3f}s/{v['iterations']} it" for k, v in cadence["trials"].items()))
    del eng
    torch.cuda.synchronize()
    # host to host: the box form, the engine (upload, K^T and SELL build), the solve, the extraction
    t0 = time.perf_counter()
    f2 = box_form(p)
    t1 = time.perf_counter()
    eng = SingleEngine(f2, "fp32", tune=tune)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    r = eng.solve(tol, mi)
    x, obj, y = extract(f2, p, r["x"], r["y"])
    t3 = time.perf_counter()
    h2h = {"s": t3 - t0, "lps_per_s": (r["status"] == "optimal") / (t3 - t0), "box_form_s": t1 - t0,
           "engine_setup_s": t2 - t1, "solve_and_extract_s": t3 - t2}
    log(f"transport {sources}x{sinks} host to host {h2h['s']:.3f}s (box form {t1 - t0:.3f}, engine {t2 - t1:.3f}, "
        f"solve + extract {t3 - t2:.3f})")
    steps = args.c1_steps
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    results = [eng.solve(tol, mi) for _ in range(steps)]
    torch.cuda.synchronize()
    secs = time.perf_counter() - t4
    ok = sum(rr["status"] == "optimal" for rr in results)
    log(f"transport resident: {ok}/{steps} optimal, {secs / steps * 1e3:.0f} ms per solve, "
        f"iterations {results[-1]['iterations']}, rel KKT {results[-1]['rel_kkt']:.2e}")
    obj_ref = highs_objective(p) if not args.no_highs else None
    us_p = eng.time_kernel("primal", eng.tune, reps=50)
    us_d = eng.time_kernel("dual", eng.tune, reps=50)
    bp = primal_bytes(form, 4, int(eng.t_sell_ptr[-1].item()) if eng.tune.sell_primal else 0,
                      values=not (eng.k_const and eng.tune.sell_primal))
    one = 2 <= eng.tune.sell <= 7
    bd = dual_bytes(form, 4, values=not (eng.k_const and one), index_loads=eng.dual_index_loads if one else None)
    gp, gd = bp / (us_p * 1e-6) / 1e9, bd / (us_d * 1e-6) / 1e9
    spmv = spmv_rates(eng, traffic_key=f"transport_{sources}x{sinks}")
    log(f"primal step {us_p:.2f} us ({gp:.0f} GB/s), dual step {us_d:.2f} us ({gd:.0f} GB/s); SpMV "
        + ", ".join(f"{k} {v['achieved_gbs']:.0f}" for k, v in spmv.items()) + " GB/s")
    key = f"transport_{sources}x{sinks}"
    rec = {
        "workload": f"C1 single transportation LP {sources}x{sinks} (m={form.m}, n={form.n}, nnz={form.nnz}), "
                    f"plain restarted PDHG from a cold start, tol {tol:g}",
        "clock": "resident: LP in HBM, one solve to tolerance per step; host_to_host: box form, engine set-up "
                 "(upload, device K^T, SELL slices), solve and extraction",
        "value": ok / secs, "unit": "LPs/s", "steps": steps, "ms_per_step": secs * 1e3 / steps,
        "iterations": results[-1]["iterations"], "restarts": results[-1]["restarts"],
        "rel_kkt": results[-1]["rel_kkt"], "objective": float(obj),
        "objective_rel_err_vs_fp64": None if obj_ref is None else abs(obj - obj_ref) / (1 + abs(obj_ref)),
        "host_to_host": h2h, "tune": tune.__dict__, "cadence": cadence,
        "roofline": {"bound": "hbm",
                     "kernel": pkern(eng) + (" (SELL-64 copy of K^T" if eng.tune.sell_primal else " (CSR of K^T")
                               + "; fused K^T y + step + projection + sums)",
                     "achieved": gp, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gp / HBM_PEAK_GBS,
                     "traffic": measured_traffic(f"{key}/{pkern(eng)}"), "bytes_per_launch": bp,
                     "us_per_launch": us_p,
                     "dual_kernel": {"achieved": gd, "frac": gd / HBM_PEAK_GBS, "bytes_per_launch": bd,
                                     "us_per_launch": us_d, "traffic": measured_traffic(f"{key}/dual_kernel")},
                     "spmv_standalone": spmv},
    }
    if cpu and not args.no_cpu_baseline:
        rf = reference_form(p)
        reng = SingleEngine(rf, "fp64")
        rr = reng.solve(tol, mi)
        ref_iters = rr["iterations"] if rr["status"] == "optimal" else None
        del reng
        rec["reference_form_iterations"] = ref_iters
        rec["cpu_baseline"] = transport_cpu(p, ref_iters or results[-1]["iterations"], args.cpu_sample_iters)
    del eng
    return rec


def transport_cpu(problem, iters_needed: int, sample_iters: int):
    """The CPU oracle per iteration on a bounded sample of the transportation LP,
    scaled to the iterations the reference algorithm needs (a full run takes minutes)."""
    _cpu_init()
    from oracle import pdhg_ref
    t0 = time.perf_counter()
    sf = pdhg_ref.to_standard_form(problem)
    tau, sigma = pdhg_ref.step_sizes(sf.A)
    AT = sf.A.T.tocsr()
    t_setup = time.perf_counter() - t0
    A, b, c = sf.A, sf.b, sf.c
    x = np.zeros(A.shape[1]); y = np.zeros(A.shape[0])
    xs = np.zeros_like(x); ys = np.zeros_like(y)
    t1 = time.perf_counter()
    for _ in range(sample_iters):            # the loop body of pdhg.py:209-222
        xn = np.maximum(x - tau * (c - AT @ y), 0.0)
        xb = 2.0 * xn - x
        y = y + sigma * (b - A @ xb)
        x = xn
        xs += x
        ys += y
    t_iter = (time.perf_counter() - t1) / sample_iters
    t2 = time.perf_counter()
    norm_b = max(1.0, float(np.linalg.norm(b))); norm_c = max(1.0, float(np.linalg.norm(c)))
    pdhg_ref.kkt(A, AT, b, c, norm_b, norm_c, xs / sample_iters, ys / sample_iters)
    t_kkt = time.perf_counter() - t2
    total = t_setup + iters_needed * (t_iter + t_kkt / 50.0)
    return {"value": 1.0 / total, "unit": "LPs/s", "cores": 1, "kind": "port",
            "sample": (f"oracle FP64 reference-form PDHG on the same LP: {sample_iters} iterations timed "
                       f"({t_iter * 1e3:.1f} ms/iter, KKT {t_kkt * 1e3:.1f} ms per 50 it, setup {t_setup:.1f} s), "
                       f"scaled to the {iters_needed} iterations the reference algorithm needs to 1e-4 "
                       f"(counted on the GPU in the reference form, FP64); 1 core, BLAS threads 1")}


# ---------------------------------------------------- C2: uneven batched LPs
def batch_lp_bytes(pk, out, mask=None):
    """Bytes the persistent kernel moves at minimum, summed over LPs.

    Per iteration of LP k: K and K^T once each at 6 B per entry (FP32 value, 16-bit
    index; the line pointers stay in LDS), 32 B per column (c, tau, lo, hi, x r/w, x
    partial sum r/w), 5 B per row (b, row kind).  Per check (two candidates): K and
    K^T twice more, 64 B per column and 41 B per row of FP64 sums and data."""
    nnz, m, n = pk.nnzs.astype(float), pk.ms.astype(float), pk.ns.astype(float)
    mat = 6 * nnz
    per_it = 2 * mat + 32 * n + 5 * m
    per_check = 4 * mat + 64 * n + 41 * m
    its = out["iterations"].astype(float)
    checks = out["checks"].astype(float) + 1
    tot = its * per_it + checks * per_check
    return float(np.sum(tot if mask is None else tot[mask]))


def run_batch(args, cpu=True):
    """Config 2: 4,096 uneven random LPs, two persistent launches."""
    import torch
    from sovereign_solver_amd.engine.batch import BatchEngine, PackedBatch
    from sovereign_solver_amd.lp.batch import results_of
    from sovereign_solver_amd.workloads import uneven_batch

    tol, mi = args.tol, args.max_iters
    probs = uneven_batch(args.batch, seed=0)
    # host to host: packing (K, K^T with 16-bit indices, descriptors), upload, solve, results
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pk = PackedBatch(probs)
    t1 = time.perf_counter()
    eng = BatchEngine(pk)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    out, xs, ys = eng.solve(tol, mi)
    res = results_of(eng, out, xs, ys)
    t3 = time.perf_counter()
    h2h = {"s": t3 - t0, "lps_per_s": sum(r.status == "optimal" for r in res) / (t3 - t0), "pack_s": t1 - t0,
           "upload_s": t2 - t1, "solve_and_results_s": t3 - t2}
    log(f"uneven batch host to host {h2h['s']:.2f}s (pack {t1 - t0:.2f}, upload {t2 - t1:.2f}, "
        f"solve + results {t3 - t2:.2f}) -> {h2h['lps_per_s']:.0f} LPs/s")
    tuned = None
    if not args.no_autotune:
        from sovereign_solver_amd.engine import tune as T
        t = time.perf_counter()
        tuned = T.tune_batch(eng, tol, mi, use_cache=False)
        tuned["s"] = time.perf_counter() - t
        log(f"persistent engine tuned in {tuned['s']:.1f}s: {tuned['config']} "
            f"(trials: { {k: round(v['seconds'], 3) for k, v in tuned['trials'].items()} })")
    steps = args.c2_steps
    times, outs = [], []
    for _ in range(steps):
        eng.set_warm_start(None, None, None)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        t = time.perf_counter()
        ev[0].record(eng.stream)
        eng.launch(tol, mi)
        ev[1].record(eng.stream)
        eng.stream.synchronize()
        times.append((time.perf_counter() - t, ev[0].elapsed_time(ev[1]) / 1e3))
        outs.append(np.frombuffer(eng.t["out"].cpu().numpy().tobytes(), dtype=out.dtype).copy())
    wall = sum(t[0] for t in times)
    n_ok = sum(int((o["status"] == 1).sum()) for o in outs)
    value = n_ok / wall
    log(f"uneven batch resident: {n_ok}/{steps * len(probs)} optimal, {wall / steps:.3f}s per batch -> "
        f"{value:.0f} LPs/s")
    # each part alone: the FP32 part's launch for the roofline
    eng.set_warm_start(None, None, None)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record(eng.stream)
    eng.launch(tol, mi, parts=1)
    ev[1].record(eng.stream)
    eng.launch(tol, mi, parts=2)
    ev[2].record(eng.stream)
    eng.results()
    t32, t64 = ev[0].elapsed_time(ev[1]) / 1e3, ev[1].elapsed_time(ev[2]) / 1e3
    b32 = batch_lp_bytes(pk, outs[-1], mask=~pk.resident)
    b64 = batch_lp_bytes(pk, outs[-1], mask=pk.resident)
    g32 = b32 / t32 / 1e9
    traffic = measured_traffic(f"uneven_{args.batch}/batch_kernel_fp32")
    log(f"FP32 part {t32:.3f}s ({b32 / 1e12:.3f} TB -> {g32:.0f} GB/s), FP64 part {t64:.3f}s")
    rng = np.random.default_rng(123)
    sub = rng.choice(len(probs), size=min(args.c2_plain, len(probs)), replace=False)
    plain = plain_rates([probs[k] for k in sub], tol, mi)
    log(f"plain single-instance on {plain['lps']} LPs: resident {plain['resident_lps_per_s']:.1f} LPs/s, "
        f"drop-in call {plain['drop_in_call_lps_per_s']:.1f} LPs/s")
    osub = rng.choice(len(probs), size=min(args.obj_sample, len(probs)), replace=False)
    # random LPs of this family can take HiGHS's simplex minutes: interior point, 5 s per LP
    oe = objective_errors([probs[k] for k in osub], [res[k].objective for k in osub], len(osub),
                          method="highs-ipm", time_limit=5.0, total_cap_s=60.0)
    log(f"objective vs HiGHS on {oe['lps']} LPs: max {oe['max']:.2e}, median {oe['median']:.2e}")
    rec = {
        "workload": f"C2 batched restarted PDHG, {len(probs)} LPs, m log-uniform in [50,2000], n=2.5m, density 1%, "
                    f"two concurrent persistent launches (FP32 part, FP64 part), cold start, tol {tol:g}",
        "clock": "resident: packed batch in HBM, both parts launched and finished per step; host_to_host: "
                 "packing, upload, solve and SolveResults",
        "value": value, "unit": "LPs/s", "steps": steps, "ms_per_step": wall * 1e3 / steps,
        "sum_m": pk.M, "sum_n": pk.N, "sum_nnz": pk.Z, "lps_fp32_part": int(len(probs) - pk.n_resident),
        "lps_fp64_part": int(pk.n_resident), "kernel_s_both_parts": times[-1][1],
        "iterations_mean": float(outs[-1]["iterations"].mean()), "iterations_max": int(outs[-1]["iterations"].max()),
        "host_to_host": h2h, "plain_single_gpu": plain, "tuned": tuned,
        "speedup_vs_plain": value / plain["resident_lps_per_s"] if plain["resident_lps_per_s"] else None,
        "objective_rel_err_vs_fp64": oe,
        "roofline": {"bound": "hbm", "kernel": f"batch_kernel<{eng.block}, float> (FP32 part: one pass over K and one over "
                                               "K^T per iteration, FP32 values and 16-bit indices)",
                     "achieved": g32, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": g32 / HBM_PEAK_GBS,
                     "traffic": traffic, "bytes_per_launch": b32, "us_per_launch": t32 * 1e6,
                     "traffic_rate_gbs": None if traffic is None else traffic / t32 / 1e9,
                     "fp64_part": {"s": t64, "bytes": b64, "achieved": b64 / t64 / 1e9 if t64 > 0 else None}},
    }
    if cpu and not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline(f"uneven/{len(probs)}", np.arange(len(probs)), tol, args.cpu_cap_s / 2,
                                           "LPs drawn uniformly from the uneven batch")
    del eng
    return rec


# ---------------------------------------------------- C4: set-cover node LPs
def run_nodes(args, cpu=True):
    """Config 4: 256 branch-and-bound node relaxations of a set-covering MILP,
    one shared-matrix batch warm-started from the parent (root) solution."""
    import torch
    from sovereign_solver_amd import solve_lp_pdhg
    from sovereign_solver_amd.engine.family import FamilyBatch, FamilyEngine
    from sovereign_solver_amd.lp.batch import family_warm_start
    from sovereign_solver_amd.workloads import SetCoverNodes

    tol, mi = args.tol, args.max_iters
    sc = SetCoverNodes()
    root = solve_lp_pdhg(sc.root(), tol=tol, precision="fp32")
    log(f"set cover {sc.rows}x{sc.cols}: root {root.status} it={root.iterations} obj={root.objective:.4f}")
    node_ids = np.arange(args.nodes)
    probs = [sc.node(int(k)) for k in node_ids]
    eng = FamilyEngine(FamilyBatch(probs))
    shape = eng.autotune(tol)
    ws = [(np.clip(root.x, p.lb, p.ub), root.y, root.step_scale) for p in probs]
    X0, Y0, w0 = family_warm_start(eng.fb, ws)
    out_host = torch.empty(eng.t["out_final"].shape, dtype=torch.uint8).pin_memory()
    from sovereign_solver_amd.engine.batch import OUT_DTYPE

    def solve_parent():
        eng.start(tol, X0, Y0, w0)
        eng.run(mi)
        eng.collect()
        with torch.cuda.stream(eng.stream):
            out_host.copy_(eng.t["out_final"], non_blocking=True)
        eng.stream.synchronize()
        return np.frombuffer(out_host.numpy().tobytes(), dtype=OUT_DTYPE).copy()

    def solve_cold():
        eng.start(tol)
        eng.run(mi)
        eng.collect()
        with torch.cuda.stream(eng.stream):
            out_host.copy_(eng.t["out_final"], non_blocking=True)
        eng.stream.synchronize()
        return np.frombuffer(out_host.numpy().tobytes(), dtype=OUT_DTYPE).copy()

    rates = {}
    for name, fn in (("parent_warm", solve_parent), ("cold", solve_cold)):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        steps = args.c4_steps
        ok = 0
        for _ in range(steps):
            o = fn()
            ok += int((o["status"] == 1).sum())
        dt = time.perf_counter() - t
        rates[name] = {"lps_per_s": ok / dt, "ms_per_batch": dt * 1e3 / steps,
                       "iters_mean": float(o["iterations"].mean()), "iters_max": int(o["iterations"].max())}
    log(f"nodes: parent warm {rates['parent_warm']}, cold {rates['cold']}")
    torch.cuda.synchronize()
    t = time.perf_counter()
    eng.load(probs)
    eng.start(tol, X0, Y0, w0)
    eng.run(mi)
    o, X, Y = eng.results()
    obj = eng.objectives(X)
    dt = time.perf_counter() - t
    h2h = {"lps_per_s": int((o["status"] == 1).sum()) / dt, "ms_per_batch": dt * 1e3}
    roof = family_passes(eng, tol, f"setcover_{len(probs)}")
    del eng
    plain = plain_rates(probs[: args.c4_plain], tol, mi)
    oe = objective_errors(probs, obj, args.c4_obj_sample, total_cap_s=60.0)
    log(f"nodes plain {plain}; objective vs HiGHS on {oe['lps']}: max {oe['max']:.2e} ({oe['s']:.1f}s)")
    rec = {
        "workload": f"C4 branch-and-bound node relaxations, set cover {sc.rows}x{sc.cols} (nnz {sc.A.nnz}), "
                    f"{len(probs)} nodes with 1-50 fixings, shared-matrix batch warm-started from the parent (root) "
                    f"solution, tol {tol:g}",
        "clock": "resident: node data in HBM, start from the parent point (host arrays), solve, results in HBM; "
                 "host_to_host adds the node data's packing and the results' copy and objectives",
        "value": rates["parent_warm"]["lps_per_s"], "unit": "LPs/s", "steps": args.c4_steps,
        "root_iterations": root.iterations, "parent_warm": rates["parent_warm"], "cold": rates["cold"],
        "host_to_host": h2h, "launch_shape": shape, "plain_single_gpu": plain,
        "speedup_vs_plain": (rates["parent_warm"]["lps_per_s"] / plain["resident_lps_per_s"]
                             if plain["resident_lps_per_s"] else None),
        "objective_rel_err_vs_fp64": oe, "roofline": roof,
    }
    if cpu and not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline("setcover", node_ids, tol, args.cpu_cap_s / 2, "the set-cover nodes")
    return rec


# ------------------------------------------------ beyond the Infinity Cache
def run_beyond_mall(args, sources=4000, sinks=8000):
    """Standalone SpMV and the fused steps on a transportation LP whose K and K^T
    (64M entries, 0.5 GB each in FP32 with int32 indices) exceed the 256 MB MALL."""
    import torch
    from sovereign_solver_amd.engine.forms import box_form
    from sovereign_solver_amd.engine.single import SingleEngine
    from sovereign_solver_amd.engine import tune as T
    from sovereign_solver_amd.workloads import transportation_lp

    t0 = time.perf_counter()
    p = transportation_lp(sources, sinks, seed=0)
    form = box_form(p)
    del p
    eng = SingleEngine(form, "fp32")
    tune = T.autotune(eng, reps=10) if not args.no_autotune else eng.tune
    eng.set_tune(tune)
    log(f"transport {sources}x{sinks} (n={form.n}, nnz={form.nnz}) built and tuned in {time.perf_counter() - t0:.1f}s")
    us_p = eng.time_kernel("primal", eng.tune, reps=10)
    us_d = eng.time_kernel("dual", eng.tune, reps=10)
    bp = primal_bytes(form, 4, int(eng.t_sell_ptr[-1].item()) if eng.tune.sell_primal else 0,
                      values=not (eng.k_const and eng.tune.sell_primal))
    one = 2 <= eng.tune.sell <= 7
    bd = dual_bytes(form, 4, values=not (eng.k_const and one), index_loads=eng.dual_index_loads if one else None)
    gp, gd = bp / (us_p * 1e-6) / 1e9, bd / (us_d * 1e-6) / 1e9
    key = f"transport_{sources}x{sinks}"
    spmv = spmv_rates(eng, reps=10, kx_tpr=64, traffic_key=key)
    log(f"beyond MALL: primal step {us_p:.1f} us ({gp:.0f} GB/s), dual step {us_d:.1f} us ({gd:.0f} GB/s), SpMV "
        + ", ".join(f"{k} {v['achieved_gbs']:.0f}" for k, v in spmv.items()) + " GB/s")
    rec = {
        "workload": f"transportation LP {sources}x{sinks} (m={form.m}, n={form.n}, nnz={form.nnz}): K and K^T "
                    f"0.5 GB each, working set of a step ~{bp / 1e9:.2f} GB > 256 MB Infinity Cache",
        "clock": "HIP graphs of back-to-back launches at the cold-start state, events on the engine stream",
        "tune": tune.__dict__,
        "primal_kernel": {"kernel": pkern(eng), "achieved": gp, "frac": gp / HBM_PEAK_GBS, "bytes_per_launch": bp,
                          "us_per_launch": us_p, "traffic": measured_traffic(f"{key}/{pkern(eng)}")},
        "dual_kernel": {"achieved": gd, "frac": gd / HBM_PEAK_GBS, "bytes_per_launch": bd, "us_per_launch": us_d,
                        "traffic": measured_traffic(f"{key}/dual_kernel")},
        "spmv_standalone": spmv,
    }
    del eng
    torch.cuda.empty_cache()
    return rec


# ---------------------------------------------------- branch-and-bound demo
def bench_bnb(args):
    """Branch-and-bound of a set-covering MILP with batched node relaxations."""
    from scipy.optimize import Bounds, LinearConstraint, milp
    from sovereign_solver_amd.milp import solve_milp_batched
    from sovereign_solver_amd.workloads import unit_cost_set_cover

    p = unit_cost_set_cover(120, 300, 0.04, seed=1)
    log(f"set cover MILP {p.n_constraints}x{p.n_vars}, nnz {p.A_ub.nnz}, unit-ish costs")
    t0 = time.perf_counter()
    ref = milp(p.c, constraints=[LinearConstraint(p.A_ub, -np.inf, p.b_ub)], integrality=np.ones(p.n_vars),
               bounds=Bounds(p.lb, p.ub), options={"time_limit": 300})
    t_highs = time.perf_counter() - t0
    log(f"HiGHS MILP (CPU): status {ref.status} objective {ref.fun:.6f} in {t_highs:.1f}s")
    solve_milp_batched(p, batch=args.bnb_batch, node_limit=2 * args.bnb_batch)          # warm-up
    t0 = time.perf_counter()
    res = solve_milp_batched(p, batch=args.bnb_batch, time_limit=300)
    t_b = time.perf_counter() - t0
    rate = res.nodes / t_b
    log(f"batched B&B (batch {args.bnb_batch}): {res.status} objective {res.objective:.6f} nodes {res.nodes} "
        f"in {t_b:.1f}s -> {rate:.0f} node LPs/s, gap {res.gap:.2e}")
    t0 = time.perf_counter()
    seq = solve_milp_batched(p, batch=1, time_limit=args.bnb_seq_s)
    t_s = time.perf_counter() - t0
    seq_rate = seq.nodes / t_s
    log(f"one node at a time (batch 1, {t_s:.0f}s sample): {seq.nodes} nodes -> {seq_rate:.0f} node LPs/s")
    out = {
        "metric": METRIC, "value": rate, "unit": "LPs/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": t_b * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 iterates, fp64 KKT", "data": "synthetic set-covering MILP",
        "config": {"workload": f"branch-and-bound, set cover {p.n_constraints}x{p.n_vars} with near-unit costs, "
                               f"best-bound rounds of {args.bnb_batch} node relaxations solved as one shared-matrix "
                               f"batch (parent warm starts), node LPs to {args.tol:g}",
                   "status": res.status, "objective": res.objective, "nodes": res.nodes, "seconds": t_b,
                   "gap": res.gap, "highs_milp_objective": float(ref.fun), "highs_milp_status": int(ref.status),
                   "highs_milp_seconds": t_highs,
                   "objective_rel_err": abs(res.objective - ref.fun) / (1 + abs(ref.fun)),
                   "one_node_at_a_time_node_lps_per_s": seq_rate,
                   "speedup_vs_one_at_a_time": rate / seq_rate if seq_rate > 0 else None},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


# ----------------------------------------------------------------------- main
def line(fields, roof, cpu, args, extra_config=None):
    cfg = dict(fields["config"])
    if extra_config:
        cfg.update(extra_config)
    return {"metric": METRIC, "value": fields["value"], "unit": "LPs/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": fields["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": fields["dtype"], "data": fields["data"],
            "config": cfg, "roofline": roof, "cpu_baseline": cpu}


def sub_as_line(rec, args):
    """A sub-record as the JSON line of a single-workload run."""
    return {"metric": METRIC, "value": rec["value"], "unit": "LPs/s", "n_gpus": 1, "steps": rec.get("steps", 1),
            "warmup": 1, "ms_per_step": rec.get("ms_per_step"), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32 iterates and SpMV, fp64 KKT", "data": "synthetic",
            "config": {k: v for k, v in rec.items() if k not in ("roofline", "cpu_baseline")},
            "roofline": rec.get("roofline"), "cpu_baseline": rec.get("cpu_baseline")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed steps of the headline workload")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "learned", "transport", "batch", "nodes", "beyond_mall", "bnb"])
    ap.add_argument("--no-subrecords", action="store_true", help="headline only (profiling runs)")
    ap.add_argument("--skip", default="", help="comma list of sub-records to leave out: c1,c2,c4,beyond_mall")
    ap.add_argument("--train", type=int, default=50_000)
    ap.add_argument("--eval", type=int, default=1024)
    ap.add_argument("--serve-batches", type=int, default=4,
                    help="held-out batches of the serving loop (host to host, overlapped)")
    ap.add_argument("--stages", type=int, default=3, help="node-model stages")
    ap.add_argument("--node-steps", type=int, default=60_000, help="training steps per stage")
    ap.add_argument("--node-rows", type=int, default=32768, help="node rows per training step")
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--obj-sample", type=int, default=64)
    ap.add_argument("--cpu-cap-s", type=float, default=12.0)
    ap.add_argument("--c1-steps", type=int, default=3)
    ap.add_argument("--c2-steps", type=int, default=2)
    ap.add_argument("--c2-plain", type=int, default=32)
    ap.add_argument("--c4-steps", type=int, default=3)
    ap.add_argument("--c4-plain", type=int, default=16)
    ap.add_argument("--c4-obj-sample", type=int, default=64)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--nodes", type=int, default=256)
    ap.add_argument("--bnb-batch", type=int, default=256, help="bnb: node relaxations per round")
    ap.add_argument("--bnb-seq-s", type=float, default=15.0, help="bnb: seconds of the one-node-at-a-time sample")
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--max-iters", type=int, default=100_000, help="per LP (the reference's MAX_PDHG_ITERS)")
    ap.add_argument("--cpu-sample-iters", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="use the heuristic launch shape")
    ap.add_argument("--no-highs", action="store_true", help="skip the HiGHS objective reference of C1")
    ap.add_argument("--calib-copy", action="store_true",
                    help="end with 1 GiB device copies (calibration of the HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); every box here has one GPU, "
              "and multi-GPU sharding of this path is design-only.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "bnb":
        return bench_bnb(args)
    single = {"transport": run_transport, "batch": run_batch, "nodes": run_nodes, "beyond_mall": run_beyond_mall}
    if args.workload in single:
        rec = single[args.workload](args)
        calib_copy(args)
        print(json.dumps(sub_as_line(rec, args) if "value" in rec else rec), flush=True)
        return
    t0 = time.perf_counter()
    fields, roof, cpu = run_headline(args)
    subs = {}
    skip = set(s.strip() for s in args.skip.split(",") if s.strip())
    if not args.no_subrecords and args.workload == "headline":
        for name, fn in (("c1", run_transport), ("c2", run_batch), ("c4", run_nodes),
                         ("beyond_mall", run_beyond_mall)):
            if name in skip:
                continue
            t = time.perf_counter()
            try:
                subs[name] = fn(args)
            except Exception as e:                       # a sub-record must not lose the headline
                log(f"sub-record {name} failed: {type(e).__name__}: {e}")
                subs[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"sub-record {name} done in {time.perf_counter() - t:.1f}s")
    log(f"bench done in {time.perf_counter() - t0:.1f}s")
    calib_copy(args)
    print(json.dumps(line(fields, roof, cpu, args, {"subrecords": subs} if subs else None)), flush=True)


if __name__ == "__main__":
    main()
