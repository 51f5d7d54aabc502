"""ctypes binding of libpdlp_hip.so (declared in include/pdlp_hip.h).

The library is built in-tree (``sovereign_solver_amd/_lib/libpdlp_hip.so``).
There is no fallback: if it is missing or fails to load, :func:`lib` raises.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "_lib", "libpdlp_hip.so")

PDLP_ABI_VERSION = 15
PDLP_F32, PDLP_F64 = 0, 1
PDLP_E_ARG, PDLP_E_TUNE, PDLP_E_SHAPE, PDLP_E_LIMIT = -1, -2, -3, -4
PDLP_RUNNING, PDLP_OPTIMAL, PDLP_NUMERICAL, PDLP_INFEASIBLE, PDLP_UNBOUNDED, PDLP_SKIPPED = 0, 1, 3, 4, 5, 6
PDLP_ROW_EQ, PDLP_ROW_LE = 0, 1
PDLP_PARTIAL_BLOCKS = 1024
PDLP_PARTIALS_LEN = 4 * PDLP_PARTIAL_BLOCKS * 16
PDLP_SCHEME_REFERENCE, PDLP_SCHEME_ADAPTIVE = 0, 1
PDLP_FAM_TILE = 64
PDLP_FAM_SHARED = {"c": 1, "b": 2, "lo": 4, "hi": 8, "c64": 16, "b64": 32, "b_shift": 64, "lo_shift": 128,
                   "ub_span": 256}
PDLP_FAM_KKT_CHUNKS = 128
PDLP_FAM_LONG = 64
PDLP_FAM_NSUM = 20

_p = C.c_void_p
_i32p = C.c_void_p


class PdlpCtrl(C.Structure):
    _fields_ = [
        ("iter", C.c_int32), ("window", C.c_int32), ("status", C.c_int32), ("restarts", C.c_int32),
        ("done_iter", C.c_int32), ("restart_pending", C.c_int32), ("apply_window", C.c_int32),
        ("checks", C.c_int32),
        ("ref_err", C.c_double), ("last_err", C.c_double), ("rp", C.c_double), ("rd", C.c_double),
        ("gap", C.c_double), ("tol", C.c_double), ("norm_b", C.c_double), ("norm_c", C.c_double),
        ("tau_scale", C.c_double), ("sigma_scale", C.c_double), ("p_obj", C.c_double), ("d_obj", C.c_double),
        ("err_prev", C.c_double), ("cand", C.c_int32), ("last_restart", C.c_int32), ("weight0", C.c_double),
    ]


class PdlpDev(C.Structure):
    _fields_ = [
        ("m", C.c_int32), ("n", C.c_int32), ("nnz", C.c_int32), ("prec", C.c_int32),
        ("row_ptr", _p), ("col_idx", _p), ("val", _p),
        ("t_row_ptr", _p), ("t_col_idx", _p), ("t_val", _p),
        ("c", _p), ("lo", _p), ("hi", _p), ("b", _p), ("row_kind", _p),
        ("c64", _p), ("b64", _p), ("b_shift", _p), ("lo_shift", _p), ("ub_span", _p),
        ("x", _p), ("xbar", _p), ("y", _p), ("x_sum", _p), ("y_sum", _p),
        ("x_psum", _p), ("y_psum", _p), ("partials", _p), ("ctrl", _p), ("x_last", _p), ("y_last", _p),
        ("sell_slices", C.c_int32), ("n_csr_rows", C.c_int32), ("sell_ptr", _p), ("sell_rows", _p),
        ("sell_col", _p), ("sell_val", _p), ("csr_rows", _p), ("sell_max_width", C.c_int32),
        ("k_const", C.c_int32), ("sell_part", _p), ("sell_asum", _p),
        ("row_scale", _p), ("col_scale", _p),
        ("t_sell_slices", C.c_int32), ("t_sell_reserved", C.c_int32), ("t_sell_ptr", _p), ("t_sell_idx", _p),
        ("t_sell_val", _p), ("sell_stride", _p), ("csr_stride", _p),
    ]


class PdlpTune(C.Structure):
    _fields_ = [
        ("tpr_primal", C.c_int32), ("tpr_dual", C.c_int32), ("block", C.c_int32),
        ("check_every", C.c_int32), ("xcd_remap", C.c_int32), ("exact", C.c_int32), ("block_dual", C.c_int32),
        ("scheme", C.c_int32), ("sell", C.c_int32), ("no_ray", C.c_int32), ("sell_primal", C.c_int32),
    ]


class PdlpBatch(C.Structure):
    _fields_ = [
        ("count", C.c_int32), ("check_every", C.c_int32), ("max_iters", C.c_int32), ("block", C.c_int32),
        ("lds_bytes", C.c_int32), ("grid", C.c_int32), ("n_resident", C.c_int32), ("block_resident", C.c_int32),
        ("lds_resident", C.c_int32), ("grid_resident", C.c_int32), ("parts", C.c_int32), ("x_in_lds", C.c_int32),
        ("lds_need", C.c_int32), ("lds_need_resident", C.c_int32), ("lines", C.c_int32),
        ("tol", C.c_double),
        ("lps", C.c_void_p), ("order", C.c_void_p), ("queue", C.c_void_p),
        ("row_ptr", C.c_void_p), ("col_idx", C.c_void_p), ("val", C.c_void_p),
        ("t_col_ptr", C.c_void_p), ("t_row_idx", C.c_void_p), ("t_val", C.c_void_p),
        ("c", C.c_void_p), ("lo", C.c_void_p), ("hi", C.c_void_p),
        ("c64", C.c_void_p), ("lo_shift", C.c_void_p), ("ub_span", C.c_void_p),
        ("b", C.c_void_p), ("row_kind", C.c_void_p), ("b64", C.c_void_p), ("b_shift", C.c_void_p),
        ("x", C.c_void_p), ("x_psum", C.c_void_p), ("x_sum", C.c_void_p), ("tau", C.c_void_p),
        ("y", C.c_void_p), ("y_sum", C.c_void_p), ("out", C.c_void_p),
        ("x_last", C.c_void_p), ("y_last", C.c_void_p), ("y_psum", C.c_void_p),
        ("row_scale", C.c_void_p), ("col_scale", C.c_void_p),
    ]


class PdlpFamilyLP(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("iter", C.c_int32), ("window", C.c_int32), ("restarts", C.c_int32),
        ("checks", C.c_int32), ("done_iter", C.c_int32), ("last_restart", C.c_int32), ("action", C.c_int32),
        ("norm_b", C.c_double), ("norm_c", C.c_double), ("ref_err", C.c_double), ("err_prev", C.c_double),
        ("w_div", C.c_double), ("weight", C.c_float), ("weight0", C.c_float),
        ("weight_cold", C.c_float), ("warm", C.c_int32), ("iter0", C.c_int32), ("reserved", C.c_int32),
        ("start_err", C.c_double), ("cold_err", C.c_double),
    ]


class PdlpFamily(C.Structure):
    _fields_ = [
        ("m", C.c_int32), ("n", C.c_int32), ("count", C.c_int32), ("tiles", C.c_int32),
        ("check_every", C.c_int32), ("flags", C.c_int32), ("shared", C.c_int32), ("items_per_wave", C.c_int32),
        ("long_rows", C.c_int32), ("long_cols", C.c_int32), ("xcd_affinity", C.c_int32),
        ("failover_iters", C.c_int32), ("max_iters", C.c_int32), ("reserved", C.c_int32),
        ("tol", C.c_double),
        ("row_ptr", _p), ("col_idx", _p), ("val", _p), ("t_row_ptr", _p), ("t_col_idx", _p), ("t_val", _p),
        ("row_kind", _p), ("tau_hat", _p), ("sigma_hat", _p), ("row_metric", _p), ("row_orig", _p),
        ("col_orig", _p), ("c", _p), ("lo", _p), ("hi", _p), ("b", _p),
        ("c64", _p), ("lo_shift", _p), ("ub_span", _p), ("b64", _p), ("b_shift", _p),
        ("x", _p), ("xbar", _p), ("x_psum", _p), ("x_last", _p),
        ("y", _p), ("y_psum", _p), ("y_last", _p), ("x_sum", _p), ("y_sum", _p),
        ("lp", _p), ("out", _p), ("run_mask", _p), ("running", _p), ("partials", _p),
    ]


class PdlpGemm(C.Structure):
    _fields_ = [
        ("M", C.c_int32), ("N", C.c_int32), ("K", C.c_int32),
        ("lda", C.c_int32), ("ldb", C.c_int32), ("ldc", C.c_int32),
        ("trans_a", C.c_int32), ("trans_b", C.c_int32), ("relu", C.c_int32), ("accumulate", C.c_int32),
        ("alpha", C.c_float), ("split_k", C.c_int32),
        ("A", _p), ("B", _p), ("bias", _p), ("mask", _p), ("C", _p), ("Cb", _p), ("work", _p),
    ]


class PdlpMlp(C.Structure):
    _fields_ = [
        ("batch", C.c_int32), ("d_in", C.c_int32), ("d_h", C.c_int32), ("d_out", C.c_int32),
        ("x", _p), ("w1", _p), ("w2", _p), ("w3", _p), ("b1", _p), ("b2", _p), ("b3", _p),
        ("h1", _p), ("h2", _p), ("out", _p),
        ("target", _p), ("out_weight", _p), ("dout", _p), ("dout_b", _p), ("dz2", _p), ("dz1", _p),
        ("gw1", _p), ("gw2", _p), ("gw3", _p), ("gb1", _p), ("gb2", _p), ("gb3", _p), ("loss", _p),
        ("work", _p), ("work_floats", C.c_int64), ("w2b", _p), ("w3b", _p),
    ]


class PdlpMlpInfer(C.Structure):
    _fields_ = [
        ("rows", C.c_int32), ("d_in", C.c_int32), ("d_h", C.c_int32), ("d_out", C.c_int32),
        ("x", _p), ("w1p", _p), ("w2p", _p), ("w3p", _p), ("b1", _p), ("b2", _p), ("b3", _p), ("out", _p),
        ("h1", _p), ("h2", _p),
    ]


class PdlpNodeGather(C.Structure):
    _fields_ = [
        ("batch", C.c_int32), ("nodes", C.c_int32), ("width", C.c_int32), ("ld_out", C.c_int32),
        ("src_len", C.c_int32), ("shared_len", C.c_int32), ("src2_len", C.c_int64), ("count", C.c_int64),
        ("seed", C.c_uint64), ("src", _p), ("shared", _p), ("src2", _p), ("table", _p), ("scale", _p), ("mean", _p), ("inv_std", _p), ("rows", _p),
        ("out_bf16", _p), ("out_f32", _p),
    ]


class PdlpNodeAssembly(C.Structure):
    _fields_ = [
        ("batch", C.c_int32), ("nodes", C.c_int32), ("ld_net", C.c_int32),
        ("len", C.c_int32), ("ld_out", C.c_int32), ("nnz", C.c_int32),
        ("indptr", _p), ("node", _p), ("chan", _p), ("weight", _p), ("bias", _p), ("net", _p), ("out", _p),
    ]


class PdlpDispatchInputs(C.Structure):
    _fields_ = [
        ("count", C.c_int32), ("G", C.c_int32), ("T", C.c_int32),
        ("ld_c", C.c_int32), ("ld_b", C.c_int32), ("ld_out", C.c_int32),
        ("off_d", C.c_int32), ("off_c", C.c_int32), ("off_u", C.c_int32), ("off_cc", C.c_int32),
        ("off_cc2", C.c_int32), ("off_p", C.c_int32), ("off_m", C.c_int32), ("off_x", C.c_int32),
        ("pmax_sum", C.c_double),
        ("c", _p), ("b", _p), ("pmax", _p), ("cost", _p), ("demand", _p), ("out", _p),
    ]


class PdlpError(RuntimeError):
    pass


_lock = threading.Lock()
_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "pdlp_abi_version": (C.c_int, []),
    "pdlp_last_error": (C.c_char_p, []),
    "pdlp_csr_transpose_work_bytes": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_size_t)]),
    "pdlp_csr_transpose": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, _p, _p, _p, _p, _p, _p, _p,
                                     C.c_size_t, _p]),
    "pdlp_spmv_csr": (C.c_int, [C.c_int32, _p, _p, _p, _p, _p, C.c_int32, C.c_int32, _p]),
    "pdlp_csr_to_sell64": (C.c_int, [C.c_int32, C.c_int32, _p, _p, _p, _p, _p, _p, _p]),
    "pdlp_sell64_widths": (C.c_int, [C.c_int32, _p, _p, _p]),
    "pdlp_row_strides": (C.c_int, [C.c_int32, _p, _p, _p, _p, _p]),
    "pdlp_arena_transpose_work_bytes": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                                  C.POINTER(C.c_size_t)]),
    "pdlp_arena_transpose": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _p, _p, _p, _p, _p,
                                       _p, _p, _p, _p, _p, _p, _p, C.c_size_t, _p]),
    "pdlp_spmv_sell64": (C.c_int, [C.c_int32, _p, _p, _p, _p, _p, C.c_int32, _p]),
    "pdlp_spmv_csr_f64acc": (C.c_int, [_p, _p, _p, _p, _p, C.c_int32, C.c_int32, _p]),
    "pdlp_init": (C.c_int, [C.POINTER(PdlpDev), C.POINTER(PdlpTune), _p]),
    "pdlp_primal_step": (C.c_int, [C.POINTER(PdlpDev), C.POINTER(PdlpTune), _p]),
    "pdlp_dual_step": (C.c_int, [C.POINTER(PdlpDev), C.POINTER(PdlpTune), _p]),
    "pdlp_kkt_check": (C.c_int, [C.POINTER(PdlpDev), C.POINTER(PdlpTune), _p]),
    "pdlp_run": (C.c_int, [C.POINTER(PdlpDev), C.POINTER(PdlpTune), C.c_int32, C.c_int32, _p]),
    "pdlp_spmv_dev": (C.c_int, [C.POINTER(PdlpDev), C.POINTER(PdlpTune), _p, _p, _p]),
    "pdlp_run_noalloc": (C.c_int, [C.POINTER(PdlpDev), C.POINTER(PdlpTune), C.c_int32, C.c_int32, _p]),
    "pdlp_graph_capture": (C.c_int, [C.POINTER(PdlpDev), C.POINTER(PdlpTune), C.c_int32, _p,
                                     C.POINTER(C.c_void_p)]),
    "pdlp_graph_launch": (C.c_int, [_p, _p]),
    "pdlp_graph_destroy": (C.c_int, [_p]),
    "pdlp_graph_capture_step": (C.c_int, [C.POINTER(PdlpDev), C.POINTER(PdlpTune), C.c_int32, C.c_int32, _p,
                                          C.POINTER(C.c_void_p)]),
    "pdlp_batch_solve": (C.c_int, [C.POINTER(PdlpBatch), _p]),
    "pdlp_family_init": (C.c_int, [C.POINTER(PdlpFamily), _p]),
    "pdlp_family_records": (C.c_int, [C.POINTER(PdlpFamily), _p, C.c_int32, _p]),
    "pdlp_family_load": (C.c_int, [C.POINTER(PdlpFamily), _p, _p, _p, _p, _p, _p]),
    "pdlp_family_run": (C.c_int, [C.POINTER(PdlpFamily), C.c_int32, _p]),
    "pdlp_family_iterate": (C.c_int, [C.POINTER(PdlpFamily), C.c_int32, _p]),
    "pdlp_family_finish": (C.c_int, [C.POINTER(PdlpFamily), C.c_int32, _p]),
    "pdlp_family_set_start": (C.c_int, [C.POINTER(PdlpFamily), _p, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                        _p, _p, _p]),
    "pdlp_family_graph_capture": (C.c_int, [C.POINTER(PdlpFamily), C.c_int32, _p, C.POINTER(C.c_void_p)]),
    "pdlp_family_graph_capture_step": (C.c_int, [C.POINTER(PdlpFamily), C.c_int32, C.c_int32, _p,
                                                 C.POINTER(C.c_void_p)]),
    "pdlp_gemm_bf16": (C.c_int, [C.POINTER(PdlpGemm), _p]),
    "pdlp_mlp_fwd_bf16": (C.c_int, [C.POINTER(PdlpMlp), _p]),
    "pdlp_mlp_bwd_bf16": (C.c_int, [C.POINTER(PdlpMlp), _p]),
    "pdlp_mlp_infer_pack": (C.c_int, [_p, _p, _p, C.c_int32, C.c_int32, C.c_int32, _p, _p, _p, _p]),
    "pdlp_mlp_dz_pack": (C.c_int, [_p, _p, C.c_int32, C.c_int32, _p, _p, _p]),
    "pdlp_mlp_infer_bf16": (C.c_int, [C.POINTER(PdlpMlpInfer), _p]),
    "pdlp_adam_step": (C.c_int, [_p, _p, _p, _p, _p, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_float,
                                 C.c_int32, _p]),
    "pdlp_cast_bf16": (C.c_int, [_p, _p, C.c_int64, _p]),
    "pdlp_node_gather": (C.c_int, [C.POINTER(PdlpNodeGather), _p]),
    "pdlp_node_assemble": (C.c_int, [C.POINTER(PdlpNodeAssembly), _p]),
    "pdlp_dispatch_inputs": (C.c_int, [C.POINTER(PdlpDispatchInputs), _p]),
}



def _bind(L, name, sig):
    fn = getattr(L, name)
    fn.restype, fn.argtypes = sig


def lib():
    """The loaded library; raises PdlpError if it is missing or stale."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise PdlpError(f"{LIB_PATH} is missing: build it with "
                            "`python -m sovereign_solver_amd.csrc.build` (this engine has no CPU path)")
        L = C.CDLL(LIB_PATH)
        for name, sig in _SIGS.items():
            _bind(L, name, sig)
        v = L.pdlp_abi_version()
        if v != PDLP_ABI_VERSION:
            raise PdlpError(f"libpdlp_hip.so ABI {v} != expected {PDLP_ABI_VERSION}; rebuild it")
        _lib = L
        return L


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().pdlp_last_error().decode(errors="replace")
        raise PdlpError(f"{what or 'pdlp call'} failed ({rc}): {msg}")


def exported_symbols() -> list:
    return sorted(_SIGS.keys())
