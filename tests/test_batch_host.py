"""Host-side packing of the batched engine (no GPU)."""

import numpy as np
import pytest
import scipy.sparse as sp

from sovereign_solver_amd.engine.batch import (LP_DTYPE, OUT_DTYPE, PackedBatch, lanes_per_row, lds_bytes_of,
                                               lds_bytes_resident, lds_bytes_x_in_lds)
from sovereign_solver_amd.lp.batch import pack_warm_start
from sovereign_solver_amd.workloads import random_sparse_lp, transportation_lp, uneven_batch, uneven_sizes


def test_descriptor_layout_matches_header():
    # PdlpBatchLP: 6 int32, 5 int64, 4 double; PdlpBatchOut: 4 int32, 7 double
    assert LP_DTYPE.itemsize == 6 * 4 + 5 * 8 + 4 * 8
    assert LP_DTYPE.fields["tpc"][1] == 16 and LP_DTYPE.fields["row_off"][1] == 24
    assert LP_DTYPE.fields["cp_off"][1] == 56 and LP_DTYPE.fields["norm_b"][1] == 64
    assert OUT_DTYPE.itemsize == 4 * 4 + 7 * 8
    assert OUT_DTYPE.fields["rel_kkt"][1] == 16


def test_packing_offsets_and_contents():
    probs = [transportation_lp(3, 4, seed=0), random_sparse_lp(5, 9, 0.3, seed=1), random_sparse_lp(7, 6, 0.4, seed=2)]
    pk = PackedBatch(probs, transpose="host")
    assert PackedBatch(probs).t_row_idx is None            # the default leaves K^T to the device
    for k in ("row_off", "col_off", "nz_off"):
        assert pk.arena[k][-1] == {"row_off": pk.M, "col_off": pk.N, "nz_off": pk.Z}[k]
    assert pk.M == sum(p.n_constraints for p in probs) and pk.N == sum(p.n_vars for p in probs)
    for k, p in enumerate(probs):
        d = pk.desc[k]
        m, n = d["m"], d["n"]
        rp = pk.row_ptr[d["rp_off"]: d["rp_off"] + m + 1]
        assert rp[0] == 0 and rp[-1] == pk.nnzs[k]
        ci = pk.col_idx[d["nnz_off"]: d["nnz_off"] + rp[-1]]
        va = pk.val[d["nnz_off"]: d["nnz_off"] + rp[-1]]
        dense = np.zeros((m, n))
        for i in range(m):
            dense[i, ci[rp[i]:rp[i + 1]]] = va[rp[i]:rp[i + 1]]
        full = np.vstack([p.A_ub.toarray(), p.A_eq.toarray()])
        assert np.array_equal(dense, full)
        # K^T by rows: column pointers at cp_off, entries at the same nnz_off
        cp = pk.t_col_ptr[d["cp_off"]: d["cp_off"] + n + 1]
        assert cp[0] == 0 and cp[-1] == pk.nnzs[k]
        ri = pk.t_row_idx[d["nnz_off"]: d["nnz_off"] + cp[-1]]
        tv = pk.t_val[d["nnz_off"]: d["nnz_off"] + cp[-1]]
        dense_t = np.zeros((n, m))
        for j in range(n):
            assert np.all(np.diff(ri[cp[j]:cp[j + 1]]) > 0)      # rows ascending within a column
            dense_t[j, ri[cp[j]:cp[j + 1]]] = tv[cp[j]:cp[j + 1]]
        assert np.array_equal(dense_t, full.T)
        assert np.array_equal(pk.c[d["col_off"]: d["col_off"] + n], p.c)
        kinds = pk.kind[d["row_off"]: d["row_off"] + m]
        assert list(kinds) == [1] * p.n_ineq + [0] * p.n_eq
    # small LPs: all in the FP64 part, whose LDS covers x, xbar, y and the line pointers
    assert pk.n_resident == len(probs) and pk.resident.all()
    assert pk.lds_resident >= max(12 * p.n_constraints + 20 * p.n_vars + 8 for p in probs)
    assert list(pk.order) == list(np.argsort(-pk.nnzs, kind="stable"))


def test_fp32_and_fp64_parts():
    # LPs over 16 KB of FP64-part LDS run in the FP32 part, which comes first in the work order
    probs = [random_sparse_lp(1100, 2750, 0.002, seed=1), random_sparse_lp(60, 150, 0.05, seed=2),
             random_sparse_lp(1300, 3250, 0.002, seed=3), random_sparse_lp(200, 500, 0.02, seed=4)]
    pk = PackedBatch(probs)
    assert list(pk.resident) == [False, True, False, True] and pk.n_resident == 2
    assert list(pk.order) == [2, 0, 3, 1]              # FP32 part, then FP64 part, each largest first
    # the FP32 part fits with x, its partial sum, tau_hat and c in LDS as well
    assert pk.x_in_lds and pk.lds_bytes == (int(lds_bytes_x_in_lds(1300, 3250)) + 15) // 16 * 16
    assert pk.lds_resident == (int(lds_bytes_resident(200, 500)) + 15) // 16 * 16
    assert pk.lds_bytes <= 160 * 1024 and pk.lds_resident <= 16 * 1024
    # an FP32-part LP beyond that LDS keeps the whole part on the smaller layout
    pk2 = PackedBatch(probs + [random_sparse_lp(400, 6000, 0.002, seed=6)])
    assert not pk2.x_in_lds and pk2.lds_bytes == (int(lds_bytes_of(400, 6000)) + 15) // 16 * 16


def test_uneven_family_statistics():
    ms = uneven_sizes(4096)
    assert ms.min() >= 50 and ms.max() <= 2000
    # log-uniform: median near sqrt(50*2000) = 316
    assert 280 < np.median(ms) < 360
    b = uneven_batch(8, seed=1)
    for p in b:
        m, n = p.A_eq.shape
        assert n == round(2.5 * m)
        assert abs(p.A_eq.nnz - round(0.01 * m * n)) <= 1


def test_lp_beyond_lds_is_rejected():
    import pytest
    big = random_sparse_lp(200, 14_000, 0.001, seed=5)       # 12 n > 160 KB
    with pytest.raises(ValueError, match="LDS"):
        PackedBatch([random_sparse_lp(5, 9, 0.3, seed=1), big])


def test_lanes_per_row():
    # lanes x 8 entries cover mean + 2 sqrt(mean) entries
    assert lanes_per_row(1.2) == 1 and lanes_per_row(4) == 1 and lanes_per_row(4.5) == 2
    assert lanes_per_row(17) == 4 and lanes_per_row(42) == 8 and lanes_per_row(50) == 16
    assert lanes_per_row(10_000) == 64


def test_warm_start_packing_clips_and_projects():
    probs = [random_sparse_lp(4, 6, 0.5, seed=3), transportation_lp(2, 3, seed=1)]
    pk = PackedBatch(probs)
    ws = [(np.full(6, 100.0), np.ones(4), 2.0), (np.full(6, -1.0), np.ones(5))]
    x0, y0, w = pack_warm_start(pk, ws)
    assert np.allclose(x0[:6], probs[0].ub)            # clipped to ub
    assert np.all(x0[6:] == 0.0)                        # clipped to lb = 0
    assert np.all(y0[4:6] == 0.0) and np.all(y0[6:] == 1.0)   # <= rows projected to y <= 0
    assert w[0] == 2.0 and np.isnan(w[1])                # NaN = the engine default


def test_scaled_batch_is_the_same_lps():
    """A scaled arena holds D_r K D_c and the matching b, c and bounds; the
    normalisers are the original LPs'; the transposes are of the scaled matrices."""
    from sovereign_solver_amd.engine import scaling as S
    probs = [random_sparse_lp(30 + 7 * k, 80 + 11 * k, 0.08, seed=k, family=2) for k in range(5)]
    plain = PackedBatch(probs)
    sc = PackedBatch(probs, scaling="ruiz_pc", transpose="host")
    np.testing.assert_array_equal(plain.desc["norm_b"], sc.desc["norm_b"])
    np.testing.assert_array_equal(plain.desc["norm_c"], sc.desc["norm_c"])
    for k, p in enumerate(probs):
        r0, c0 = sc.row_off[k], sc.col_off[k]
        m, n = sc.ms[k], sc.ns[k]
        dr, dc = sc.row_scale[r0: r0 + m], sc.col_scale[c0: c0 + n]
        K0 = sp.vstack([p.A_ub, p.A_eq]).tocsr()
        Dr, Dc = S.scales(K0, "ruiz_pc")
        np.testing.assert_allclose(dr, Dr)
        np.testing.assert_allclose(dc, Dc)
        z0, rp0 = sc.nnz_off[k], int(sc.desc["rp_off"][k])
        K = sp.csr_matrix((sc.val[z0: z0 + K0.nnz], sc.col_idx[z0: z0 + K0.nnz].astype(np.int64),
                           sc.row_ptr[rp0: rp0 + m + 1]), shape=(m, n))
        np.testing.assert_allclose(K.toarray(), (sp.diags(Dr) @ K0 @ sp.diags(Dc)).toarray(), rtol=1e-12)
        cp0 = sc.cp_off[k]
        KT = sp.csr_matrix((sc.t_val[z0: z0 + K0.nnz], sc.t_row_idx[z0: z0 + K0.nnz].astype(np.int64),
                            sc.t_col_ptr[cp0: cp0 + n + 1]), shape=(n, m))
        np.testing.assert_allclose(KT.toarray(), K.toarray().T, rtol=1e-12)
        np.testing.assert_allclose(sc.c[c0: c0 + n], p.c * Dc)
        np.testing.assert_allclose(sc.lo[c0: c0 + n], p.lb / Dc)
        np.testing.assert_allclose(sc.hi[c0: c0 + n], p.ub / Dc)
        b = np.concatenate([p.b_ub, p.b_eq])
        np.testing.assert_allclose(sc.b[r0: r0 + m], b * Dr)
        np.testing.assert_allclose(sc.b_shift[r0: r0 + m], plain.b_shift[r0: r0 + m] * Dr)
    assert plain.row_scale is None and plain.col_scale is None
    with pytest.raises(ValueError):
        PackedBatch(probs, scaling="bogus")
