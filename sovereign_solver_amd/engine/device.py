"""Device plumbing shared by the engines: torch tensors as HIP buffers."""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import PdlpError


def require_gpu() -> torch.device:
    """The HIP device; raises when none is visible (there is no CPU path)."""
    if not torch.cuda.is_available():
        raise PdlpError("no HIP device visible: the PDHG engine runs only on the GPU")
    return torch.device("cuda", torch.cuda.current_device())


def dev_tensor(a, dtype, device) -> torch.Tensor:
    """Upload a host array (contiguous) as a device tensor of ``dtype``."""
    t = torch.from_numpy(np.ascontiguousarray(a))
    return t.to(device=device, dtype=dtype, non_blocking=False).contiguous()


def ptr(t) -> int:
    return 0 if t is None else t.data_ptr()


def stream_handle(s: torch.cuda.Stream) -> C.c_void_p:
    return C.c_void_p(s.cuda_stream)


def struct_to_tensor(st: C.Structure, device) -> torch.Tensor:
    raw = bytes(C.string_at(C.addressof(st), C.sizeof(st)))
    host = torch.frombuffer(bytearray(raw), dtype=torch.uint8)
    return host.to(device)


def struct_from_bytes(cls, t: torch.Tensor):
    """Decode a host uint8 tensor holding one ctypes Structure."""
    return cls.from_buffer_copy(t.numpy().tobytes())


def csr_transpose(row_ptr: torch.Tensor, col_idx: torch.Tensor, val: torch.Tensor, cols: int,
                  stream: torch.cuda.Stream):
    """K^T of a device CSR matrix (rows x cols) on the device, through the C ABI's
    ``pdlp_csr_transpose``: ``(t_row_ptr, t_col_idx, t_val)``, SciPy's canonical
    transpose bit for bit.  Enqueued on ``stream``."""
    from . import _lib as L
    rows, nnz = row_ptr.numel() - 1, col_idx.numel()
    if val.dtype not in (torch.float32, torch.float64) or val.numel() != nnz:
        raise ValueError("val must be float32 or float64 with one value per entry")
    prec = L.PDLP_F32 if val.dtype == torch.float32 else L.PDLP_F64
    lib = L.lib()
    need = C.c_size_t(0)
    L.check(lib.pdlp_csr_transpose_work_bytes(rows, cols, nnz, C.byref(need)), "pdlp_csr_transpose_work_bytes")
    dv = row_ptr.device
    with torch.cuda.stream(stream):
        t_row_ptr = torch.empty(cols + 1, dtype=torch.int32, device=dv)
        t_col_idx = torch.empty(nnz, dtype=torch.int32, device=dv)
        t_val = torch.empty(nnz, dtype=val.dtype, device=dv)
        work = torch.empty(max(int(need.value), 1), dtype=torch.uint8, device=dv)
        L.check(lib.pdlp_csr_transpose(prec, rows, cols, nnz, ptr(row_ptr), ptr(col_idx), ptr(val), ptr(t_row_ptr),
                                       ptr(t_col_idx), ptr(t_val), ptr(work), int(need.value),
                                       stream_handle(stream)), "pdlp_csr_transpose")
        work.record_stream(stream)
    return t_row_ptr, t_col_idx, t_val
