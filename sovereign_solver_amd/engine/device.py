"""Device plumbing shared by the engines: torch tensors as HIP buffers."""

from __future__ import annotations

import ctypes as C
import threading
from typing import Optional

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


_NP_OF = {torch.float32: np.float32, torch.float64: np.float64, torch.int32: np.int32, torch.int8: np.int8,
          torch.int16: np.int16, torch.uint8: np.uint8, torch.int64: np.int64}
_STAGING_KEEP = 512 << 20          # pinned staging kept between uploads up to this size


class _Staging:
    """One pinned host buffer reused by uploads (grown on demand), and the event of
    the last copies out of it."""
    lock = threading.Lock()
    buf: Optional[torch.Tensor] = None
    ev: Optional[torch.cuda.Event] = None


def upload(items, device, stream: torch.cuda.Stream) -> list:
    """Device tensors of host arrays ``items`` = [(array or None, torch dtype), ...]: each
    converted to its dtype while copied into one pinned staging buffer, then copied
    to the device asynchronously on ``stream`` (None stays None).  The staging buffer
    is kept for the next upload (up to 512 MB); an upload first waits for the
    previous one's copies out of it."""
    arrs, offs, total = [], [], 0
    for a, dt in items:
        if a is None:
            arrs.append(None)
            offs.append(0)
            continue
        a = np.asarray(a).reshape(-1)
        arrs.append((a, dt))
        offs.append(total)
        total += (a.size * torch.empty(0, dtype=dt).element_size() + 255) // 256 * 256
    out = []
    with _Staging.lock:
        if _Staging.ev is not None:
            _Staging.ev.synchronize()
        buf = _Staging.buf
        if buf is None or buf.numel() < total:
            buf = torch.empty(max(total, 1 << 20), dtype=torch.uint8).pin_memory()
        host = buf.numpy()
        with torch.cuda.stream(stream):
            for it, off in zip(arrs, offs):
                if it is None:
                    out.append(None)
                    continue
                a, dt = it
                nb = a.size * torch.empty(0, dtype=dt).element_size()
                np.copyto(host[off: off + nb].view(_NP_OF[dt]), a, casting="unsafe")
                d = torch.empty(a.size, dtype=dt, device=device)
                if a.size:
                    d.copy_(buf[off: off + nb].view(dt), non_blocking=True)
                out.append(d)
            ev = torch.cuda.Event()
            ev.record(stream)
        if buf.numel() <= _STAGING_KEEP:
            _Staging.buf, _Staging.ev = buf, ev
        else:                              # too large to keep: released once its copies are done
            ev.synchronize()
            _Staging.ev = None
    return out


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
