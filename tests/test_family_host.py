"""Host side of the shared-matrix (family) engine: tiled layout, shared-vector
detection, per-LP normalisers and the reference-form yardstick data (no GPU)."""

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import yardstick
from sovereign_solver_amd.engine import _lib as L
from sovereign_solver_amd.engine.family import FamilyBatch, from_tiles, to_tiles
from sovereign_solver_amd.engine.forms import box_form
from sovereign_solver_amd.problem import Problem
from sovereign_solver_amd.workloads import DispatchFamily, SetCoverNodes, random_sparse_lp


def test_tiles_round_trip_and_address_formula():
    rng = np.random.default_rng(0)
    for B, Ln in ((1, 3), (64, 5), (65, 7), (200, 2)):
        a = rng.standard_normal((B, Ln))
        tiles = (B + 63) // 64
        flat = to_tiles(a, tiles)
        assert flat.size == tiles * 64 * Ln
        np.testing.assert_array_equal(from_tiles(flat, tiles, Ln, B), a)
        # element d of LP k sits at ((k // 64) * L + d) * 64 + k % 64
        for k in (0, B - 1, B // 2):
            for d in (0, Ln - 1):
                assert flat[((k // 64) * Ln + d) * 64 + k % 64] == a[k, d]


def test_dispatch_family_shares_bounds_not_costs():
    fam = DispatchFamily(G=10, T=4)
    fb = FamilyBatch([fam.problem(k) for k in range(5)])
    assert fb.is_shared("lo") and fb.is_shared("hi") and fb.is_shared("lo_shift") and fb.is_shared("ub_span")
    assert not fb.is_shared("c") and not fb.is_shared("b") and not fb.is_shared("b_shift")
    assert fb.tiles == 1 and fb.K.shape == (fam.m, fam.n)


def test_setcover_nodes_share_costs_not_bounds():
    sc = SetCoverNodes(rows=60, cols=300, density=0.03)
    fb = FamilyBatch([sc.node(k) for k in range(4)])
    assert fb.is_shared("c") and fb.is_shared("b") and fb.is_shared("c64")
    assert not fb.is_shared("lo") and not fb.is_shared("hi") and not fb.is_shared("b_shift")


def test_shared_copies_match_the_problems():
    """The single copies of shared vectors are the problems' own vectors in device
    order, with the lifted-form shift; step sizes are the Pock-Chambolle diagonal ones."""
    sc = SetCoverNodes(rows=60, cols=300, density=0.03)
    probs = [sc.node(k) for k in range(3)]
    fb = FamilyBatch(probs)
    np.testing.assert_array_equal(fb.from_device_cols(fb.shared_vec["c"][None])[0], probs[0].c)
    np.testing.assert_array_equal(fb.from_device_rows(fb.shared_vec["b"][None])[0], probs[0].b_ub)
    assert "lo" not in fb.shared_vec and "b_shift" not in fb.shared_vec
    fam = DispatchFamily(G=10, T=4)
    fd = FamilyBatch([fam.problem(k) for k in range(5)])
    f = box_form(fam.problem(0))
    np.testing.assert_array_equal(fd.shared_vec["lo_shift"], np.zeros(fd.n))
    np.testing.assert_array_equal(fd.from_device_cols(fd.shared_vec["ub_span"][None])[0], f.ub_span)
    colsum = np.asarray(abs(fb.K).sum(axis=0)).ravel()
    np.testing.assert_allclose(fb.tau_hat, 0.999 / np.maximum(colsum, 1e-4), rtol=1e-7)


def test_with_problems_checks_family_and_sharing():
    fam = DispatchFamily(G=10, T=4)
    fb = FamilyBatch([fam.problem(k) for k in range(5)])
    fb2 = fb.with_problems([fam.problem(k) for k in range(5, 10)])
    assert fb2.count == 5 and fb2.K is fb.K and fb2.problems[0].name == "dispatch_5"
    with pytest.raises(ValueError, match="LPs for an engine"):
        fb.with_problems([fam.problem(k) for k in range(4)])
    other = fam.problem(3)
    other = Problem(c=other.c, A_ub=other.A_ub, b_ub=other.b_ub, A_eq=other.A_eq, b_eq=other.b_eq,
                    lb=other.lb, ub=other.ub * 2, integer_mask=other.integer_mask)
    with pytest.raises(ValueError, match="share"):
        fb.with_problems([fam.problem(k) for k in range(4)] + [other])


def test_rejects_different_matrices():
    a, b = random_sparse_lp(20, 40, 0.1, seed=1), random_sparse_lp(20, 40, 0.1, seed=2)
    with pytest.raises(ValueError, match="matrix differs"):
        FamilyBatch([a, b])


def test_equal_but_distinct_matrix_objects_accepted():
    p = random_sparse_lp(20, 40, 0.1, seed=1)
    q = Problem(c=p.c * 2, A_ub=p.A_ub.copy(), b_ub=p.b_ub, A_eq=p.A_eq.copy(), b_eq=p.b_eq, lb=p.lb, ub=p.ub,
                integer_mask=p.integer_mask)
    fb = FamilyBatch([p, q])
    assert fb.count == 2 and not fb.is_shared("c")


def test_struct_layouts():
    import ctypes as C
    # 8 int32, 5 doubles, weight, weight0, weight_cold, warm, iter0, reserved, start_err, cold_err
    assert C.sizeof(L.PdlpFamilyLP) == 8 * 4 + 5 * 8 + 4 * 6 + 8 + 8
    assert C.sizeof(L.PdlpFamily) == 14 * 4 + 8 + 35 * 8
    lib = L.lib()
    assert lib.pdlp_family_init(None, None) == -1                               # PDLP_E_ARG
    assert b"NULL" in lib.pdlp_last_error()
    f = L.PdlpFamily(m=3, n=4, count=65, tiles=1, check_every=50, items_per_wave=4)
    assert lib.pdlp_family_run(C.byref(f), 1, None) == -3                    # PDLP_E_SHAPE: tiles


def test_compaction_index_moves_slots():
    """The flat gather index of a slot permutation maps tile layout to tile layout."""
    import torch
    from sovereign_solver_amd.engine.family import FamilyEngine
    rng = np.random.default_rng(3)
    B, Ln = 150, 7
    tiles = 3
    a = rng.standard_normal((B, Ln))
    flat = torch.from_numpy(to_tiles(a, tiles))
    keep = np.sort(rng.choice(B, size=70, replace=False))
    new_T = 2
    perm = np.concatenate([keep, np.full(new_T * 64 - keep.size, keep[0])])
    idx = FamilyEngine._slot_index(None, torch.from_numpy(perm), Ln, new_T)
    moved = torch.take(flat, idx).numpy()
    np.testing.assert_array_equal(from_tiles(moved, new_T, Ln, keep.size), a[keep])


def test_device_numbering_puts_long_rows_last():
    """Rows / columns above PDLP_FAM_LONG entries are numbered last on the device,
    and the permuted K is the problem's K with rows and columns renumbered."""
    fam = DispatchFamily(G=80, T=4)                  # demand rows have 80 > 64 entries
    fb = FamilyBatch([fam.problem(k) for k in range(3)])
    assert fb.long_rows == fam.T and fb.long_cols == 0
    assert np.all(np.diff(fb.K.indptr)[-fb.long_rows:] > L.PDLP_FAM_LONG)
    assert np.all(np.diff(fb.K.indptr)[: fb.m - fb.long_rows] <= L.PDLP_FAM_LONG)
    dense = fb.K_orig.toarray()
    np.testing.assert_array_equal(fb.K.toarray(), dense[fb.row_perm][:, fb.col_perm])
    rng = np.random.default_rng(0)
    X = rng.standard_normal((3, fb.n))
    np.testing.assert_array_equal(fb.from_device_cols(fb.to_device_cols(X)), X)
    Y = rng.standard_normal((3, fb.m))
    np.testing.assert_array_equal(fb.from_device_rows(fb.to_device_rows(Y)), Y)
    # device-order shared data matches the problems' own data
    np.testing.assert_array_equal(fb.from_device_cols(fb.shared_vec["hi"][None])[0], fam.problem(1).ub)
