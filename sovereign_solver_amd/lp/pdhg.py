"""Drop-in for the reference's ``solve_lp_pdhg`` (``solver/lp/pdhg.py:152-259``).

Same call and result: ``solve_lp_pdhg(problem, max_iters=MAX_PDHG_ITERS,
tol=PDHG_TOL) -> SolveResult`` with ``status in {"optimal",
"iteration_limit"}``, ``x`` in the original variable space clipped to the
bounds, the objective in the original sense and the iteration count.  The
work runs in HIP kernels on the GPU (there is no CPU path: without a HIP
device or the built library the call raises).

Keyword-only extensions (defaults keep the reference's contract):

``formulation``
    ``"box"`` (default): the engine's native form, bounds by projection,
    ``<=`` rows by a sign-constrained dual; ``"reference"``: the reference's
    slack/upper-bound-row standard form, for trajectory-level parity.
``precision``
    ``"fp32"`` iterates and SpMV with FP64 sums, KKT and objective, or
    ``"fp64"`` throughout; ``"auto"`` (default) picks fp32 when
    ``tol >= 1e-5`` and fp64 below, since FP32 iterates stall near 1e-6.
``exact``
    With ``formulation="reference", precision="fp64"``: one lane per row,
    ascending summation and no FMA contraction, i.e. SciPy's ``csr_matvec``
    arithmetic; the iterates then match the reference's bit for bit up to
    the KKT reductions (whose order differs at the 1e-16 level).
``scheme``
    ``"reference"``: the reference's restart rule (window average, restart when
    the KKT falls below half its value at the last restart) with fixed steps;
    ``"adaptive"``: restart candidate = better of average and current iterate,
    restarts on sufficient/necessary decay or a long window, and a primal
    weight updated at each restart from the distances moved since the last
    one (``"adaptive_start"``: from the distances to the starting point;
    ``"adaptive_fixed"``: weight kept); ``"auto"`` (default): adaptive for the
    box form, reference for ``formulation="reference"`` and ``exact``.
``x0, y0, step_scale``
    Warm start (box form: x in the original space, y ordered ``[A_ub; A_eq]``)
    and the starting primal weight ω: τ ← ω·τ, σ ← σ/ω (ω = 1 is the
    reference's step; None = 1 for the reference scheme, a norm-ratio
    estimate for the adaptive one).
``check_every``, ``tune``
    KKT/restart cadence (50 = reference) and launch shape (default: the
    autotuner's heuristic).  ``tune="auto"``: launch shape and check cadence
    from the autotuner (:func:`~sovereign_solver_amd.engine.tune.autotune`
    times the kernels; :func:`~sovereign_solver_amd.engine.tune.tune_cadence`
    solves this LP once per candidate cadence and keeps the fastest optimal
    solve), cached per shape bucket, so the cost is paid once per bucket.
``scaling``
    Diagonal preconditioning of the box form (``"none"`` default,
    ``"ruiz"``, ``"geometric"``, ``"pock_chambolle"``, ``"ruiz_pc"``; the
    reference's ``solver/utils/scaling.py`` scalers, which its PDHG never
    calls).  PDHG iterates on the scaled LP; the tolerance applies to the
    KKT yardstick of the original LP, and x, y come back unscaled.  The
    returned ``step_scale`` is the primal weight of the scaled LP.

The reference returns no duals and no residuals; the result additionally
carries ``y``, ``rel_kkt``, ``rp``, ``rd``, ``gap`` and ``restarts``.

Status ``"unbounded"`` (not with the reference scheme): the LP has a ray of
decrease, i.e. it is dual infeasible — unbounded when it is feasible.  The
estimate d = x − (window average) is tested at every check of the adaptive
schemes: c'd < 0 and ``|viol(Kd)|(1+|y|) + |viol_bounds(d)|(1+|c−K'y|) <=
(−c'd)/M`` with the violations of K d against the rows' recession cone
(``=`` rows: 0, ``<=`` rows: ≤ 0) and of d against the bounds'.  A dual
solution (y, r) bounds c'd below by ``−|y||viol(Kd)| − |r||viol_bounds(d)|``,
so a bounded LP fails the test while the iterate's duals are within a factor
M of a dual solution in norm: M = 1e3 when the rows' violation is below 1e-3
of the row sums of |K_ij d_j| (K d cancels), else 1e6.  The same bound with a
dual *solution* gives the second test, which does not depend on the iterate:
``|viol(Kd)| + |viol_bounds(d)| <= 1e-6 (−c'd)`` — a bounded LP passes it only
when every dual solution has ``|y*| + |r*| >= 1e6``, so an LP whose optimum is
merely far away (a row with coefficients of 1e-5, say) keeps iterating.  The
host then re-checks the ray in FP64 on the caller's data (:func:`ray_is_certified`,
``2e-6``); a ray that fails there (the device tested it against the iterate
precision's copy of K) is dropped, and the solve resumes from where it stopped
with the unboundedness test off.  The result then has ``x=None``, objective
``-inf`` (the reference simplex's convention) and ``ray`` = d in the original
space with ``max |d_j| = 1``.

Status ``"infeasible"`` (not with the reference scheme, which keeps the
reference's behaviour of running to the iteration limit): the dual ``y``
(the current iterate or the window average, whichever passed) is a Farkas
certificate, ``b_shift'y - max over the box of (K'y)'(x - lb_f) > 0`` with
``y <= 0`` on the ``<=`` rows, checked in FP64 with a relative margin of
1e-6 so a feasible LP is never flagged.  A certificate that needs ``K'y = 0``
exactly on a column unbounded in the needed direction (free columns) is
not found: such LPs end at the iteration limit.  ``"numerical_error"``: a
NaN or Inf in a KKT evaluation.
"""

from __future__ import annotations

from typing import Optional

import numpy as np

from ..config import MAX_PDHG_ITERS, PDHG_TOL
from ..engine.forms import extract, make_form
from ..engine.single import SingleEngine
from ..engine.tune import Tune, heuristic_tune, tune_cadence
from ..engine.tune import autotune as autotune_shape
from ..problem import Problem
from ..result import SolveResult


def _precision(precision: str, tol: float) -> str:
    if precision == "auto":
        return "fp32" if tol >= 1e-5 else "fp64"
    if precision not in ("fp32", "fp64"):
        raise ValueError("precision must be 'auto', 'fp32' or 'fp64'")
    return precision


def _scheme(scheme: str, formulation: str, exact: bool) -> int:
    if scheme == "auto":
        return 0 if (exact or formulation == "reference") else 1
    codes = {"reference": 0, "adaptive": 1, "adaptive_fixed": 2, "adaptive_start": 3}
    if scheme not in codes:
        raise ValueError(f"scheme must be 'auto' or one of {sorted(codes)}")
    if exact and scheme != "reference":
        raise ValueError("exact=True reproduces the reference scheme only")
    return codes[scheme]


def _warm_start(form, p: Problem, x0, y0):
    if x0 is None and y0 is None:
        return None, None
    if form.kind != "box":
        raise ValueError("warm starts are supported with formulation='box'")
    x = None if x0 is None else np.clip(np.asarray(x0, float), p.lb, p.ub)
    y = None
    if y0 is not None:
        y = np.asarray(y0, float).copy()
        if form.m_ub:
            y[: form.m_ub] = np.minimum(y[: form.m_ub], 0.0)
    if form.col_scale is not None:                      # into the scaled LP's space
        x = None if x is None else x / form.col_scale
        y = None if y is None else y / form.row_scale
    return x, y


RAY_EPS_HOST = 2e-6


def _ray_in_original_space(form, p: Problem, ray) -> np.ndarray:
    d = np.asarray(ray[: p.n_vars], float)
    if form.col_scale is not None:
        d = d * form.col_scale
    m = np.abs(d).max() if d.size else 0.0
    return d / m if m > 0 else d


def ray_quality(p: Problem, d: np.ndarray):
    """``(c'd, |viol(K d)|_2, |viol_bounds(d)|_2)`` in FP64 on the caller's data: the
    rows' violations are ``A_eq d`` and ``max(A_ub d, 0)``, the bounds' ``min(d_j, 0)``
    where lb_j is finite and ``max(d_j, 0)`` where ub_j is finite."""
    d = np.asarray(d, float)
    viol = []
    if p.n_ineq:
        viol.append(np.maximum(p.A_ub @ d, 0.0))
    if p.n_eq:
        viol.append(p.A_eq @ d)
    vr = float(np.linalg.norm(np.concatenate(viol))) if viol else 0.0
    lb, ub = np.asarray(p.lb, float), np.asarray(p.ub, float)
    bnd = np.where(np.isfinite(lb), np.minimum(d, 0.0), 0.0) + np.where(np.isfinite(ub), np.maximum(d, 0.0), 0.0)
    return float(np.asarray(p.c, float) @ d), vr, float(np.linalg.norm(bnd))


def ray_is_certified(p: Problem, d: np.ndarray, eps: float = RAY_EPS_HOST) -> bool:
    """A ray of decrease to within ``eps``: ``c'd < 0`` and ``|viol(K d)| +
    |viol_bounds(d)| <= eps (-c'd)``.  For a bounded LP with a dual solution (y*, r*)
    every d has ``c'd >= -|y*| |viol(K d)| - |r*| |viol_bounds(d)|``, so a bounded LP
    passes only when ``|y*| + |r*| >= 1 / eps``."""
    cd, vr, vb = ray_quality(p, d)
    return bool(np.isfinite(cd) and cd < 0.0 and vr + vb <= eps * -cd)


def solve_lp_pdhg(
    problem: Problem,
    max_iters: int = MAX_PDHG_ITERS,
    tol: float = PDHG_TOL,
    *,
    formulation: str = "box",
    precision: str = "auto",
    exact: bool = False,
    x0: Optional[np.ndarray] = None,
    y0: Optional[np.ndarray] = None,
    step_scale: Optional[float] = None,
    scheme: str = "auto",
    check_every: int = 50,
    tune=None,
    use_graphs: bool = True,
    scaling: str = "none",
) -> SolveResult:
    form = make_form(problem, formulation, scaling)
    if form.m == 0 or form.n == 0:                      # pdhg.py:171-172
        return SolveResult("optimal", np.zeros(problem.n_vars), 0.0, 0)
    prec = _precision(precision, tol)
    sch = _scheme(scheme, formulation, exact)
    auto = isinstance(tune, str)
    if auto and tune != "auto":
        raise ValueError("tune must be a Tune, None or 'auto'")
    if exact:
        tune = Tune(1, 1, 256, check_every, 0, 1)
    elif tune is None or auto:
        tune = heuristic_tune(form.m, form.n, form.nnz, check_every)
    eng = SingleEngine(form, prec, tune=tune, check_every=check_every, exact=exact, use_graphs=use_graphs,
                       scheme=sch)
    d = None
    try:
        xs, ys = _warm_start(form, problem, x0, y0)
        if auto and not exact:
            eng.set_tune(autotune_shape(eng))
            tune_cadence(eng, tol, max_iters, x0=xs, y0=ys, step_scale=step_scale)
        r = eng.solve(tol, max_iters, xs, ys, step_scale=step_scale)
        if r["status"] == "unbounded":
            d = _ray_in_original_space(form, problem, r["ray"])
            if not ray_is_certified(problem, d):
                # the device's test ran on the iterate precision's copy of K; the ray does not
                # hold in FP64 on the caller's data, so the LP is treated as not unbounded
                r = eng.resume_without_ray_test()
    finally:
        eng.drop_graphs()
    if r["status"] == "unbounded":
        # c'x (the minimised objective) decreases along d without bound: -inf, or +inf in the
        # reported sense of a 'max' problem, whose c is already negated
        return SolveResult("unbounded", None, -np.inf if problem.sense != "max" else np.inf, r["iterations"],
                           y=None, rel_kkt=r["rel_kkt"], rp=r["rp"], rd=r["rd"], gap=r["gap"],
                           restarts=r["restarts"], step_scale=r["ctrl"].tau_scale, ray=d)
    x, obj, y = extract(form, problem, r["x"], r["y"])
    return SolveResult(r["status"], x, obj, r["iterations"], y=y, rel_kkt=r["rel_kkt"], rp=r["rp"],
                       rd=r["rd"], gap=r["gap"], restarts=r["restarts"], step_scale=r["ctrl"].tau_scale)
