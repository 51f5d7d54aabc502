"""GPU: ``pdlp_run_noalloc`` (scratch in the library's static device array) against
``pdlp_run`` (the caller's workspace), through the C ABI.

Both run the same kernels on the same state, so the iterates, the FP64 window sums
and the control block must agree bit for bit.  Two LPs in a row, the second smaller,
so the second call runs on an array that still holds the first call's contents: each
call writes every scratch value it reads.  The first LP has SELL-64 slices and runs
the split SELL dual step, whose partial dot products live in the array too."""

import ctypes as C

import numpy as np
import pytest
import torch

from sovereign_solver_amd.engine import _lib as L
from sovereign_solver_amd.engine.forms import box_form
from sovereign_solver_amd.engine.single import SingleEngine
from sovereign_solver_amd.workloads import random_sparse_lp, transportation_lp

pytestmark = pytest.mark.gpu


def _engine(p, sell):
    eng = SingleEngine(box_form(p), "fp32", scheme=L.PDLP_SCHEME_ADAPTIVE)
    if sell:
        assert eng.sell_slices > 0
        eng.set_tune(eng.tune.with_(sell=8))
    return eng


def _blocks(eng, noalloc: bool, blocks: int):
    """``blocks`` x (50 iterations + check) from the same cold start; the state after."""
    lib = eng.lib
    w = eng.initial_weight()
    eng.start(1e-9, tau_scale=w, sigma_scale=1.0 / w)          # tol 1e-9: every block runs
    dev = L.PdlpDev.from_buffer_copy(eng.dev)
    if noalloc:                                                 # the library's array instead
        dev.partials = None
        dev.sell_part = None
    ct = eng.tune.to_c()
    fn = lib.pdlp_run_noalloc if noalloc else lib.pdlp_run
    for _ in range(blocks):
        L.check(fn(C.byref(dev), C.byref(ct), eng.tune.check_every, 1, eng.sh), fn.__name__)
    eng.stream.synchronize()
    return [t.clone().cpu() for t in (eng.x, eng.y, eng.x_sum, eng.y_sum, eng.ctrl)]


def test_noalloc_matches_pdlp_run_twice_in_a_row(gpu):
    lps = [(transportation_lp(60, 120, seed=2), True), (random_sparse_lp(40, 100, 0.05, seed=5), False)]
    for p, sell in lps:                                         # the second call's LP is smaller
        eng = _engine(p, sell)
        ref = _blocks(eng, False, 6)
        got = _blocks(eng, True, 6)
        for name, a, b in zip(("x", "y", "x_sum", "y_sum", "ctrl"), ref, got):
            assert torch.equal(a, b), (p.name, name)
        st = L.PdlpCtrl.from_buffer_copy(bytes(got[4].numpy()))
        assert st.iter == 300 and st.checks == 6 and np.isfinite(st.last_err), (p.name, st.iter, st.checks)
        eng.drop_graphs()


def test_noalloc_rejects_sell_scratch_beyond_its_array(gpu):
    eng = _engine(transportation_lp(60, 120, seed=2), True)
    dev = L.PdlpDev.from_buffer_copy(eng.dev)
    dev.sell_max_width = 1 << 20                                # 64 x slices x 2^17 floats: beyond 4 MB
    ct = eng.tune.to_c()
    assert eng.lib.pdlp_run_noalloc(C.byref(dev), C.byref(ct), 50, 1, eng.sh) == L.PDLP_E_LIMIT
    assert b"scratch" in eng.lib.pdlp_last_error()
    eng.drop_graphs()
