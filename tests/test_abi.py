"""The C ABI library loads and exports every entry point the headers declare
(no GPU needed: only version / error / argument-validation calls)."""

import ctypes as C
import glob
import os
import re

import pytest

from sovereign_solver_amd.engine import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(pdlp_\w+)\s*\(", txt))
    return sorted(names)


def test_library_built():
    assert os.path.exists(L.LIB_PATH), "run __graft_entry__.build() first"


def test_every_declared_symbol_exported_and_bound():
    lib = L.lib()
    decl = declared_symbols()
    assert len(decl) >= 10
    raw = C.CDLL(L.LIB_PATH)
    for name in decl:
        assert hasattr(raw, name), f"{name} declared in include/ but not exported"
    bound = set(L.exported_symbols())
    missing = [d for d in decl if d not in bound]
    assert not missing, f"declared but not bound in _lib.py: {missing}"
    assert lib.pdlp_abi_version() == L.PDLP_ABI_VERSION


def test_struct_sizes_match_header():
    # PdlpCtrl: 8 int32 + 12 double + (double, 2 int32, double); PdlpDev: 4 int32 + 27 pointers; PdlpTune: 8 int32
    assert C.sizeof(L.PdlpCtrl) == 8 * 4 + 13 * 8 + 2 * 4 + 8
    # ... then the SELL-64 block (2 int32, 5 pointers, 2 int32, 2 pointers), the two scale vectors
    # the SELL-64 copy of K^T (2 int32, 3 pointers) and the index runs (2 pointers)
    assert C.sizeof(L.PdlpDev) == 4 * 4 + 27 * 8 + 2 * 4 + 5 * 8 + 2 * 4 + 2 * 8 + 2 * 8 + 2 * 4 + 3 * 8 + 2 * 8
    assert C.sizeof(L.PdlpTune) == 11 * 4
    # PdlpGemm: 10 int32 (M..accumulate), alpha, split_k, 7 pointers (A, B, bias, mask, C, Cb, work)
    assert C.sizeof(L.PdlpGemm) == 10 * 4 + 4 + 4 + 7 * 8
    # PdlpMlp: 4 int32, 23 pointers, then the backward work space (pointer, int64), w2b, w3b
    assert C.sizeof(L.PdlpMlp) == 4 * 4 + 23 * 8 + 8 + 8 + 2 * 8
    # PdlpNodeGather: 6 int32, src2_len, count, seed, 10 pointers; PdlpNodeAssembly: 6 int32, 7 pointers
    assert C.sizeof(L.PdlpNodeGather) == 6 * 4 + 3 * 8 + 10 * 8
    assert C.sizeof(L.PdlpNodeAssembly) == 6 * 4 + 7 * 8
    # PdlpMlpInfer: 4 int32, 10 pointers
    assert C.sizeof(L.PdlpMlpInfer) == 4 * 4 + 10 * 8


def test_fused_inference_rejects_shapes_without_device():
    lib = L.lib()
    d = L.PdlpMlpInfer(rows=4, d_in=48, d_h=256, d_out=8)                # d_h must be 512
    assert lib.pdlp_mlp_infer_bf16(C.byref(d), None) == L.PDLP_E_SHAPE
    for d_in, d_out in ((44, 8), (136, 8), (48, 33), (48, 0)):
        d = L.PdlpMlpInfer(rows=4, d_in=d_in, d_h=512, d_out=d_out)
        assert lib.pdlp_mlp_infer_bf16(C.byref(d), None) == L.PDLP_E_SHAPE
    d = L.PdlpMlpInfer(rows=4, d_in=48, d_h=512, d_out=8)                 # NULL pointers
    assert lib.pdlp_mlp_infer_bf16(C.byref(d), None) == L.PDLP_E_ARG
    assert lib.pdlp_mlp_infer_pack(None, None, None, 48, 512, 8, None, None, None, None) == L.PDLP_E_ARG
    assert lib.pdlp_mlp_infer_pack(None, None, None, 48, 500, 8, None, None, None, None) == L.PDLP_E_SHAPE


def test_node_kernels_reject_bad_descriptors_without_device():
    lib = L.lib()
    assert lib.pdlp_node_gather(None, None) == L.PDLP_E_ARG
    g = L.PdlpNodeGather(batch=2, nodes=3, width=4, ld_out=2)          # ld_out < width
    assert lib.pdlp_node_gather(C.byref(g), None) == L.PDLP_E_SHAPE
    g = L.PdlpNodeGather(batch=2, nodes=3, width=4, ld_out=8)          # no output
    assert lib.pdlp_node_gather(C.byref(g), None) == L.PDLP_E_ARG
    a = L.PdlpNodeAssembly(batch=70000, nodes=1, ld_net=8, len=4, ld_out=4)
    assert lib.pdlp_node_assemble(C.byref(a), None) == L.PDLP_E_LIMIT
    a = L.PdlpNodeAssembly(batch=1, nodes=1, ld_net=8, len=4, ld_out=3)
    assert lib.pdlp_node_assemble(C.byref(a), None) == L.PDLP_E_SHAPE
    m = L.PdlpMlp(batch=4, d_in=8, d_h=8, d_out=8, work_floats=16)     # work_floats without work
    for f in ("target", "dout", "dout_b", "dz2", "dz1", "gw1", "gw2", "gw3", "gb1", "gb2", "gb3", "loss"):
        setattr(m, f, 16)                                                # any non-NULL value: rejected first
    assert lib.pdlp_mlp_bwd_bf16(C.byref(m), None) == L.PDLP_E_ARG


def test_dispatch_inputs_rejects_bad_descriptors_without_device():
    lib = L.lib()
    assert lib.pdlp_dispatch_inputs(None, None) == L.PDLP_E_ARG
    T, G = 24, 200
    ok = dict(count=4, G=G, T=T, ld_c=4800, ld_b=9224, ld_out=3 * T + 4 * G + T * G, off_d=0, off_c=T, off_u=T + G,
              off_cc=T + 2 * G, off_cc2=T + 3 * G, off_p=T + 4 * G, off_m=2 * T + 4 * G, off_x=3 * T + 4 * G,
              pmax_sum=1.0)
    a = L.PdlpDispatchInputs(**ok)                                       # NULL arrays
    assert lib.pdlp_dispatch_inputs(C.byref(a), None) == L.PDLP_E_ARG
    a = L.PdlpDispatchInputs(**dict(ok, count=0))                       # nothing to do
    assert lib.pdlp_dispatch_inputs(C.byref(a), None) == 0
    a = L.PdlpDispatchInputs(**dict(ok, G=1025, ld_c=5000))
    assert lib.pdlp_dispatch_inputs(C.byref(a), None) == L.PDLP_E_LIMIT
    a = L.PdlpDispatchInputs(**dict(ok, ld_out=ok["ld_out"] - 1))       # the X block overruns the row
    assert lib.pdlp_dispatch_inputs(C.byref(a), None) == L.PDLP_E_SHAPE
    a = L.PdlpDispatchInputs(**dict(ok, ld_b=T - 1))
    assert lib.pdlp_dispatch_inputs(C.byref(a), None) == L.PDLP_E_SHAPE
    a = L.PdlpDispatchInputs(**dict(ok, pmax_sum=0.0))
    assert lib.pdlp_dispatch_inputs(C.byref(a), None) == L.PDLP_E_ARG


def test_argument_errors_without_device():
    lib = L.lib()
    rc = lib.pdlp_run(None, None, 1, 1, None)
    assert rc == -1
    assert b"NULL" in lib.pdlp_last_error()
    t = L.PdlpTune(3, 1, 256, 50, 0, 0)
    d = L.PdlpDev()
    rc = lib.pdlp_primal_step(C.byref(d), C.byref(t), None)
    assert rc in (-1, -2, -3)
    with pytest.raises(L.PdlpError):
        L.check(rc, "primal")


def test_header_constants_match_binding():
    txt = open(os.path.join(ROOT, "include", "pdlp_hip.h")).read()
    defs = dict(re.findall(r"#define\s+(PDLP_\w+)\s+([^/\n]+)", txt))
    env = {}
    for k, v in defs.items():
        try:
            env[k] = eval(v.strip(), {}, dict(env))
        except (NameError, SyntaxError):
            pass
    for name in ("PDLP_ABI_VERSION", "PDLP_RUNNING", "PDLP_OPTIMAL", "PDLP_NUMERICAL", "PDLP_INFEASIBLE",
                 "PDLP_UNBOUNDED", "PDLP_SKIPPED", "PDLP_PARTIAL_BLOCKS", "PDLP_PARTIALS_LEN"):
        assert env[name] == getattr(L, name), name


def test_transpose_argument_errors_without_device():
    lib = L.lib()
    b = C.c_size_t(0)
    assert lib.pdlp_csr_transpose_work_bytes(-1, 2, 3, C.byref(b)) == -3          # PDLP_E_SHAPE
    assert lib.pdlp_csr_transpose_work_bytes(1, 2, 3, None) == -1                 # PDLP_E_ARG
    assert lib.pdlp_csr_transpose(7, 1, 1, 0, None, None, None, None, None, None, None, 0, None) == -1
    assert b"prec" in lib.pdlp_last_error()


def _batch_args(count=1, n_resident=0):
    """A PdlpBatch whose pointers are non-NULL host addresses: the argument checks run
    before anything touches them (no launch follows a rejected call)."""
    buf = (C.c_char * 64)()
    a = L.PdlpBatch(count=count, check_every=50, max_iters=100, block=512, lds_bytes=1024, n_resident=n_resident,
                    block_resident=256, lds_resident=1024, tol=1e-4)
    for name, typ in L.PdlpBatch._fields_:
        if typ is C.c_void_p:
            setattr(a, name, C.addressof(buf))
    return a, buf


def test_batch_rejects_lps_beyond_lds_without_device():
    lib = L.lib()
    a, _buf = _batch_args()
    a.lds_need = 2048                                   # an LP of the FP32 part needs more than lds_bytes
    assert lib.pdlp_batch_solve(C.byref(a), None) == -4                    # PDLP_E_LIMIT
    assert b"lds_bytes" in lib.pdlp_last_error()
    a, _buf = _batch_args(count=2, n_resident=1)
    a.lds_need, a.lds_need_resident = 512, 4096         # the FP64 part's LP does not fit
    assert lib.pdlp_batch_solve(C.byref(a), None) == -4
    assert b"lds_resident" in lib.pdlp_last_error()
    a, _buf = _batch_args()
    a.lds_need = 0                                      # the need must be stated
    assert lib.pdlp_batch_solve(C.byref(a), None) == -1                    # PDLP_E_ARG


def test_packer_states_each_parts_lds_need():
    from sovereign_solver_amd.engine.batch import (PackedBatch, lds_bytes_of, lds_bytes_resident,
                                                   lds_bytes_x_in_lds)
    from sovereign_solver_amd.workloads import uneven_batch
    pk = PackedBatch(uneven_batch(40, seed=3))
    fp32 = ~pk.resident
    need = (lds_bytes_x_in_lds if pk.x_in_lds else lds_bytes_of)(pk.ms, pk.ns)
    assert pk.lds_need == int(need[fp32].max()) <= pk.lds_bytes
    assert pk.lds_need_resident == int(lds_bytes_resident(pk.ms, pk.ns)[pk.resident].max()) <= pk.lds_resident


def test_noalloc_rejects_oversized_sell_scratch_without_device():
    # the size check comes before any device call
    lib = L.lib()
    d = L.PdlpDev(m=1, n=1, nnz=1, prec=L.PDLP_F32, sell_slices=4, sell_max_width=1 << 20)
    t = L.PdlpTune(1, 1, 256, 50, 0, 0, 0, 1, 8, 0, 0)
    assert lib.pdlp_run_noalloc(C.byref(d), C.byref(t), 1, 1, None) == L.PDLP_E_LIMIT
    assert b"scratch" in lib.pdlp_last_error()
    assert lib.pdlp_run_noalloc(None, C.byref(t), 1, 1, None) == L.PDLP_E_ARG


def test_sell_columns_layout_and_padding_limit():
    import numpy as np
    import scipy.sparse as sp
    from sovereign_solver_amd.engine.sell import build_columns
    # 3 rows x 130 columns: column j has j % 3 + 1 entries except the last two (empty)
    rows, cols = [], []
    for j in range(128):
        for i in range(j % 3 + 1):
            rows.append(i)
            cols.append(j)
    K = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(3, 130))
    ptr = build_columns(K, padding_max=10.0)
    # three slices (64, 64, 2 columns) of widths 3, 3, 0 entries per lane
    assert ptr.tolist() == [0, 192, 384, 384]
    assert build_columns(K, padding_max=0.5) is None              # 384 > 1.5 * 256 entries
    assert build_columns(sp.csr_matrix((3, 5))) is None


def test_arena_transpose_rejects_bad_arguments_without_device():
    lib = L.lib()
    need = C.c_size_t(0)
    assert lib.pdlp_arena_transpose_work_bytes(0, 1, 1, 1, C.byref(need)) == L.PDLP_E_SHAPE
    assert lib.pdlp_arena_transpose_work_bytes(2, 10, 12, 30, None) == L.PDLP_E_ARG
    assert lib.pdlp_arena_transpose_work_bytes(2, 10, 12, 30, C.byref(need)) == 0 and need.value > 30 * 8
    args = [L.PDLP_F32, 2, 10, 12, 30] + [None] * 11 + [None, 0, None]
    assert lib.pdlp_arena_transpose(*args) == L.PDLP_E_ARG                  # NULL arrays
    args[0] = 7
    assert lib.pdlp_arena_transpose(*args) == L.PDLP_E_ARG                  # precision


def test_csr_to_sell64_rejects_bad_arguments_without_device():
    lib = L.lib()
    assert lib.pdlp_sell64_widths(-1, None, None, None) == L.PDLP_E_SHAPE
    assert lib.pdlp_sell64_widths(5, None, None, None) == L.PDLP_E_ARG
    assert lib.pdlp_sell64_widths(0, None, None, None) == 0
    assert lib.pdlp_row_strides(-1, None, None, None, None, None) == L.PDLP_E_SHAPE
    assert lib.pdlp_row_strides(3, None, None, None, None, None) == L.PDLP_E_ARG
    assert lib.pdlp_row_strides(0, None, None, None, None, None) == 0
    assert lib.pdlp_spmv_sell64(L.PDLP_F32, None, None, None, None, None, 5, None) == L.PDLP_E_ARG
    assert lib.pdlp_spmv_sell64(L.PDLP_F32, None, None, None, None, None, -1, None) == L.PDLP_E_ARG
    assert lib.pdlp_spmv_sell64(L.PDLP_F32, None, None, None, None, None, 0, None) == 0
    assert lib.pdlp_csr_to_sell64(9, 4, None, None, None, None, None, None, None) == L.PDLP_E_ARG
    assert lib.pdlp_csr_to_sell64(L.PDLP_F32, -1, None, None, None, None, None, None, None) == L.PDLP_E_SHAPE
    assert lib.pdlp_csr_to_sell64(L.PDLP_F32, 4, None, None, None, None, None, None, None) == L.PDLP_E_ARG


def test_index_runs_of_rows_and_slices():
    """engine.sell.row_strides and SellSlices.stride: the constant column step of a row
    (of every row of a slice, at full width) or 0."""
    import numpy as np
    import scipy.sparse as sp
    from sovereign_solver_amd.engine.forms import box_form
    from sovereign_solver_amd.engine.sell import build, row_strides
    from sovereign_solver_amd.workloads import transportation_lp
    A = sp.csr_matrix(np.array([[0, 1, 0, 1, 0, 1], [1, 1, 0, 0, 0, 0], [0] * 6, [0, 0, 5, 0, 0, 0],
                                [1, 0, 1, 1, 0, 0]], float))
    assert list(row_strides(A, np.arange(5))) == [2, 1, 0, 1, 0]
    assert list(row_strides(A, np.array([4, 0]))) == [0, 2]
    K = box_form(transportation_lp(30, 150, seed=0)).K
    sl = build(K)
    assert sl.slices == 2 and np.all(sl.stride == 150)          # demand row j: columns j, j + 150, ...
    rs = row_strides(K, sl.csr_rows)
    assert np.all(rs[sl.csr_rows < 30] == 1)                     # supply rows: contiguous columns
    for i, s_ in zip(sl.csr_rows, rs):
        idx = K.indices[K.indptr[i]: K.indptr[i + 1]]
        assert s_ == 0 or np.array_equal(idx, idx[0] + s_ * np.arange(idx.size))


def test_sell_row_slices_hold_each_row_in_csr_order():
    """engine.sell.build: every row is in exactly one slice lane or in csr_rows, and a
    slice lane holds its row's entries in CSR order, then zero padding — for equal
    consecutive rows (one block of entries) and for rows of unequal length."""
    import numpy as np
    import scipy.sparse as sp
    from sovereign_solver_amd.engine.forms import box_form
    from sovereign_solver_amd.engine.sell import SLICE, build
    from sovereign_solver_amd.workloads import transportation_lp
    rng = np.random.default_rng(3)
    m = 300
    lens = rng.integers(3, 6, m)                                   # unequal rows, neighbouring columns
    rows = np.repeat(np.arange(m), lens)
    cols = np.concatenate([i + 400 * np.arange(k) for i, k in zip(range(m), lens)])
    ragged = sp.csr_matrix((rng.standard_normal(rows.size), (rows, cols)), shape=(m, 400 * 6))
    for K in (box_form(transportation_lp(100, 300, seed=1)).K, ragged):
        K = sp.csr_matrix(K)
        K.sort_indices()
        sl = build(K)
        assert sl.slices > 0
        seen = np.zeros(K.shape[0], int)
        seen[sl.csr_rows] += 1
        for s in range(sl.slices):
            w = (sl.ptr[s + 1] - sl.ptr[s]) // SLICE
            c = sl.col[sl.ptr[s]: sl.ptr[s + 1]].reshape(w, SLICE)
            v = sl.val[sl.ptr[s]: sl.ptr[s + 1]].reshape(w, SLICE)
            for lane in range(SLICE):
                r = sl.rows[s * SLICE + lane]
                if r < 0:
                    assert not v[:, lane].any()
                    continue
                seen[r] += 1
                a, b = K.indptr[r], K.indptr[r + 1]
                assert np.array_equal(c[: b - a, lane], K.indices[a:b])
                assert np.array_equal(v[: b - a, lane], K.data[a:b]) and not v[b - a:, lane].any()
        assert np.all(seen == 1)
