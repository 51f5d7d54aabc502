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
