"""Host-side formulations (no GPU): the engine's reference form equals the
oracle's restatement of pdhg.py:62-125; the box form's fused yardstick data
gives the same relative KKT as the explicit lift."""

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import pdhg_ref, yardstick
from sovereign_solver_amd.engine.forms import box_form, extract, reference_form
from sovereign_solver_amd.problem import Problem
from sovereign_solver_amd.workloads import random_sparse_lp, transportation_lp
from tests.golden.stress_instances import golden_problem


@pytest.mark.parametrize("name", ["afiro", "blending", "brandy", "adlittle", "random_n50_m30", "badly_scaled_1e8"])
def test_reference_form_matches_oracle(name):
    p = golden_problem(name)
    f = reference_form(p)
    sf = pdhg_ref.to_standard_form(p)
    assert f.K.shape == sf.A.shape
    assert np.array_equal(f.K.indptr, sf.A.indptr)
    assert np.array_equal(f.K.indices, sf.A.indices)
    assert np.array_equal(f.K.data, sf.A.data)
    assert np.array_equal(f.b, sf.b) and np.array_equal(f.c, sf.c)
    assert f.norm_b == max(1.0, float(np.linalg.norm(sf.b)))


def _fused_kkt(f, x, y):
    """The device formula (forms.py docstring) evaluated with NumPy."""
    K = f.K
    ax = K @ x
    res = ax - f.b
    if f.row_kind is not None:
        le = f.row_kind == 1
        res[le] = np.maximum(res[le], 0.0)
        rd2 = float(np.sum(np.maximum(y[le], 0.0) ** 2))
    else:
        rd2 = 0.0
    r = f.c - K.T @ y
    span = f.ub_span if f.ub_span is not None else np.full(f.n, np.inf)
    fin = np.isfinite(span)
    rneg = np.minimum(r, 0.0)
    rd2 += float(np.sum(rneg[~fin] ** 2))
    sh = f.lo_shift if f.lo_shift is not None else 0.0
    po = float(f.c @ (x - sh))
    do = float(f.b_shift @ y) + float(np.sum(span[fin] * rneg[fin]))
    rp = np.sqrt(res @ res) / f.norm_b
    rd = np.sqrt(rd2) / f.norm_c
    gap = abs(po - do) / (1 + abs(po) + abs(do))
    return max(rp, rd, gap), rp, rd, gap


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_box_form_fused_kkt_equals_explicit_lift(seed):
    rng = np.random.default_rng(seed)
    # inequality rows, equality rows, finite/inf upper bounds, nonzero lower bounds
    n, mu, me = 12, 5, 4
    p = Problem(c=rng.standard_normal(n), A_ub=sp.random(mu, n, 0.4, random_state=seed, format="csr"),
                b_ub=rng.standard_normal(mu), A_eq=sp.random(me, n, 0.4, random_state=seed + 7, format="csr"),
                b_eq=rng.standard_normal(me), lb=rng.uniform(-1, 0.5, n),
                ub=np.where(rng.random(n) < 0.5, rng.uniform(1, 3, n), np.inf), integer_mask=np.zeros(n, bool))
    f = box_form(p)
    x = np.clip(rng.standard_normal(n), p.lb, p.ub)
    y = rng.standard_normal(mu + me)
    y[:mu] = np.minimum(y[:mu], 0.0)
    got = _fused_kkt(f, x, y)
    want = yardstick.rel_kkt(p, x, y)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-14)


def test_box_form_sizes_transport():
    p = transportation_lp(10, 20, seed=0)
    f = box_form(p)
    assert f.K.shape == (30, 200) and f.K.nnz == 400
    assert f.lo is None and f.hi is None            # x >= 0, no upper bounds
    assert list(np.unique(f.row_kind)) == [0, 1]
    assert np.isclose(p.b_eq.sum(), 0.95 * p.b_ub.sum())


def test_extract_reference_form_maps_back():
    p = random_sparse_lp(8, 15, 0.3, seed=2)
    f = reference_form(p)
    z = np.abs(np.random.default_rng(0).standard_normal(f.n))
    y = np.random.default_rng(1).standard_normal(f.m)
    x, obj, yo = extract(f, p, z, y)
    assert x.shape == (15,) and yo.shape == (8,)
    assert np.all(x <= p.ub + 1e-12) and np.all(x >= p.lb - 1e-12)
    assert np.isclose(obj, p.c @ x)
