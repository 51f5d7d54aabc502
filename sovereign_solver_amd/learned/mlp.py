"""Warm-start MLP driven by the HIP kernels of libpdlp_hip.so.

Three layers d_in → d_h → d_h → d_out with ReLU; BF16 weights and
activations feed MFMA GEMMs with FP32 accumulation; FP32 master weights,
biases, gradients and Adam moments.  Torch tensors only hold the buffers;
every arithmetic step (forward, loss gradient, backward GEMMs, bias
gradients, Adam, casts) is a kernel of the library.

Parameter layout in one flat FP32 buffer: W1 [d_in][d_h], b1, W2 [d_h][d_h],
b2, W3 [d_h][d_out], b3 (weights [fan_in][fan_out]).
"""

from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np
import torch

from ..engine import _lib as L
from ..engine.device import require_gpu, stream_handle


class WarmStartMLP:
    def __init__(self, d_in: int, d_h: int, d_out: int, seed: int = 0, device=None,
                 stream: Optional[torch.cuda.Stream] = None):
        self.lib = L.lib()
        self.device = device or require_gpu()
        self.d_in, self.d_h, self.d_out = d_in, d_h, d_out
        shapes = [("W1", (d_in, d_h)), ("b1", (d_h,)), ("W2", (d_h, d_h)), ("b2", (d_h,)),
                  ("W3", (d_h, d_out)), ("b3", (d_out,))]
        self.offsets = {}
        off = 0
        for name, shp in shapes:
            self.offsets[name] = (off, shp)
            off += int(np.prod(shp))
        self.n_params = off
        rng = np.random.default_rng(np.random.SeedSequence([7, seed]))
        host = np.zeros(off, np.float32)
        for name, (o, shp) in self.offsets.items():
            if name.startswith("W"):
                host[o: o + int(np.prod(shp))] = (rng.standard_normal(int(np.prod(shp))) *
                                                  np.sqrt(2.0 / shp[0])).astype(np.float32)
        dv = self.device
        self.stream = stream if stream is not None else torch.cuda.Stream(device=dv)
        with torch.cuda.stream(self.stream):
            self.w = torch.from_numpy(host).to(dv)
            self.wb = torch.empty(off, dtype=torch.bfloat16, device=dv)
            self.g = torch.zeros(off, dtype=torch.float32, device=dv)
            self.m = torch.zeros(off, dtype=torch.float32, device=dv)
            self.v = torch.zeros(off, dtype=torch.float32, device=dv)
            self.loss = torch.zeros(1, dtype=torch.float32, device=dv)
        self.step = 0
        self._bufs = {}
        self._packed = None          # fused-inference weights, packed from the current BF16 copy
        self._packed_dz = None       # fused-backward weights, likewise
        self.fused = True            # fused kernels where the shapes allow (False: layer by layer)
        self.refresh_bf16()

    # ---------------------------------------------------------------- helpers
    @property
    def sh(self):
        return stream_handle(self.stream)

    def _p(self, name, t):
        o, _ = self.offsets[name]
        return t.data_ptr() + o * t.element_size()

    def refresh_bf16(self):
        with torch.cuda.stream(self.stream):
            L.check(self.lib.pdlp_cast_bf16(self.w.data_ptr(), self.wb.data_ptr(), self.n_params, self.sh), "cast")
        self._packed = self._packed_dz = None

    def params_numpy(self) -> dict:
        self.stream.synchronize()
        w = self.w.cpu().numpy()
        return {k: w[o: o + int(np.prod(s))].reshape(s).astype(np.float64) for k, (o, s) in self.offsets.items()}

    def set_params(self, P: dict):
        host = np.zeros(self.n_params, np.float32)
        for k, (o, s) in self.offsets.items():
            host[o: o + int(np.prod(s))] = np.asarray(P[k], np.float32).reshape(-1)
        with torch.cuda.stream(self.stream):
            self.w.copy_(torch.from_numpy(host).to(self.device))
        self.refresh_bf16()

    def grads_numpy(self) -> dict:
        self.stream.synchronize()
        g = self.g.cpu().numpy()
        return {k: g[o: o + int(np.prod(s))].reshape(s).astype(np.float64) for k, (o, s) in self.offsets.items()}

    # batches from this size on get a work space for the K-split weight gradients
    SPLIT_BATCH = 2048

    def _buffers(self, B: int, backward: bool = False):
        """Activation buffers of a B-row batch; the backward scratch is allocated on
        the first backward pass of that size (inference never holds it)."""
        dv, bf, f32 = self.device, torch.bfloat16, torch.float32
        b = self._bufs.get(B)
        if b is None:
            with torch.cuda.stream(self.stream):
                b = dict(h1=torch.empty(B, self.d_h, dtype=bf, device=dv),
                         h2=torch.empty(B, self.d_h, dtype=bf, device=dv),
                         out=torch.empty(B, self.d_out, dtype=f32, device=dv))
            self._bufs[B] = b
        if backward and "dout" not in b:
            dh, di, do = self.d_h, self.d_in, self.d_out
            wf = 32 * max(dh * dh, di * dh, dh * do) if B >= self.SPLIT_BATCH else 0
            with torch.cuda.stream(self.stream):
                b.update(dout=torch.empty(B, do, dtype=f32, device=dv),
                         dout_b=torch.empty(B, do, dtype=bf, device=dv),
                         dz2=torch.empty(B, dh, dtype=bf, device=dv),
                         dz1=torch.empty(B, dh, dtype=bf, device=dv),
                         work=torch.empty(max(wf, 1), dtype=f32, device=dv), work_floats=wf)
        return b

    def _desc(self, xb: torch.Tensor, target: Optional[torch.Tensor] = None,
              out_weight: Optional[torch.Tensor] = None) -> L.PdlpMlp:
        B = xb.shape[0]
        if xb.dtype != torch.bfloat16 or xb.shape[1] != self.d_in or not xb.is_contiguous():
            raise ValueError("input must be a contiguous [B, d_in] bfloat16 device tensor")
        bu = self._buffers(B, backward=target is not None)
        d = L.PdlpMlp(batch=B, d_in=self.d_in, d_h=self.d_h, d_out=self.d_out, x=xb.data_ptr(),
                      w1=self._p("W1", self.wb), w2=self._p("W2", self.wb), w3=self._p("W3", self.wb),
                      b1=self._p("b1", self.w), b2=self._p("b2", self.w), b3=self._p("b3", self.w),
                      h1=bu["h1"].data_ptr(), h2=bu["h2"].data_ptr(), out=bu["out"].data_ptr())
        if target is not None:
            if target.shape != (B, self.d_out) or target.dtype != torch.float32 or not target.is_contiguous():
                raise ValueError("target must be a contiguous [B, d_out] float32 device tensor")
            d.target = target.data_ptr()
            d.out_weight = 0 if out_weight is None else out_weight.data_ptr()
            d.dout, d.dout_b = bu["dout"].data_ptr(), bu["dout_b"].data_ptr()
            d.dz2, d.dz1 = bu["dz2"].data_ptr(), bu["dz1"].data_ptr()
            d.gw1, d.gw2, d.gw3 = self._p("W1", self.g), self._p("W2", self.g), self._p("W3", self.g)
            d.gb1, d.gb2, d.gb3 = self._p("b1", self.g), self._p("b2", self.g), self._p("b3", self.g)
            d.loss = self.loss.data_ptr()
            if bu["work_floats"]:
                d.work, d.work_floats = bu["work"].data_ptr(), bu["work_floats"]
            if self.fused and self.can_dz_fused:
                pb = self._pack_dz()
                d.w2b, d.w3b = pb["w2b"].data_ptr(), pb["w3b"].data_ptr()
        return d

    # ------------------------------------------------------------------- API
    def forward(self, xb: torch.Tensor) -> torch.Tensor:
        """[B, d_out] FP32 output (a view of an internal buffer, valid until the next call);
        the hidden activations stay in the internal buffers for a backward pass."""
        d = self._desc(xb)
        self._forward_desc(d)
        return self._buffers(xb.shape[0])["out"]

    def _forward_desc(self, d: L.PdlpMlp) -> None:
        """Forward pass of a descriptor: the fused kernel (storing h1 and h2) where its
        shape limits allow, else layer by layer."""
        with torch.cuda.stream(self.stream):
            if self.fused and self.can_infer_fused:
                pk = self._pack_infer()
                fi = L.PdlpMlpInfer(rows=d.batch, d_in=self.d_in, d_h=self.d_h, d_out=self.d_out, x=d.x,
                                    w1p=pk["w1p"].data_ptr(), w2p=pk["w2p"].data_ptr(), w3p=pk["w3p"].data_ptr(),
                                    b1=d.b1, b2=d.b2, b3=d.b3, out=d.out, h1=d.h1, h2=d.h2)
                L.check(self.lib.pdlp_mlp_infer_bf16(C.byref(fi), self.sh), "mlp fwd (fused)")
            else:
                L.check(self.lib.pdlp_mlp_fwd_bf16(C.byref(d), self.sh), "mlp fwd")

    def forward_layered(self, xb: torch.Tensor) -> torch.Tensor:
        """The forward pass layer by layer (three GEMM launches), whatever the shape."""
        d = self._desc(xb)
        with torch.cuda.stream(self.stream):
            L.check(self.lib.pdlp_mlp_fwd_bf16(C.byref(d), self.sh), "mlp fwd")
        return self._buffers(xb.shape[0])["out"]

    # fused inference (pdlp_mlp_infer_bf16) limits
    INFER_D_H, INFER_MAX_D_IN, INFER_MAX_D_OUT = 512, 128, 32

    @property
    def can_infer_fused(self) -> bool:
        return (self.d_h == self.INFER_D_H and self.d_in <= self.INFER_MAX_D_IN and self.d_in % 8 == 0
                and self.d_out <= self.INFER_MAX_D_OUT)

    def _pack_infer(self):
        """The BF16 weights in the fused kernel's operand-fragment order (once per update)."""
        if self._packed is None:
            k1 = -(-self.d_in // 16) * 16
            with torch.cuda.stream(self.stream):
                pk = dict(w1p=torch.empty(self.d_h * k1, dtype=torch.bfloat16, device=self.device),
                          w2p=torch.empty(self.d_h * self.d_h, dtype=torch.bfloat16, device=self.device),
                          w3p=torch.empty(32 * 32 * 16, dtype=torch.bfloat16, device=self.device))
                L.check(self.lib.pdlp_mlp_infer_pack(self._p("W1", self.wb), self._p("W2", self.wb),
                                                     self._p("W3", self.wb), self.d_in, self.d_h, self.d_out,
                                                     pk["w1p"].data_ptr(), pk["w2p"].data_ptr(),
                                                     pk["w3p"].data_ptr(), self.sh), "mlp infer pack")
            self._packed = pk
        return self._packed

    @property
    def can_dz_fused(self) -> bool:
        return self.d_h == self.INFER_D_H and self.d_out % 8 == 0 and self.d_out <= self.INFER_MAX_D_OUT

    def _pack_dz(self):
        """W2 and W3 in the fused backward kernel's operand-fragment order (once per update)."""
        if self._packed_dz is None:
            nk3 = -(-self.d_out // 16)
            with torch.cuda.stream(self.stream):
                pk = dict(w2b=torch.empty(self.d_h * self.d_h, dtype=torch.bfloat16, device=self.device),
                          w3b=torch.empty(self.d_h * 16 * nk3, dtype=torch.bfloat16, device=self.device))
                L.check(self.lib.pdlp_mlp_dz_pack(self._p("W2", self.wb), self._p("W3", self.wb), self.d_h,
                                                  self.d_out, pk["w2b"].data_ptr(), pk["w3b"].data_ptr(), self.sh),
                        "mlp dz pack")
            self._packed_dz = pk
        return self._packed_dz

    def infer(self, xb: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """[B, d_out] FP32 output of the fused inference kernel (the three layers in one
        launch; the same BF16 rounding points as :meth:`forward`, no activations kept).
        Falls back to :meth:`forward` for shapes outside the fused kernel's limits."""
        if xb.dtype != torch.bfloat16 or xb.dim() != 2 or xb.shape[1] != self.d_in or not xb.is_contiguous():
            raise ValueError("input must be a contiguous [B, d_in] bfloat16 device tensor")
        if not self.can_infer_fused:
            return self.forward(xb)
        B = xb.shape[0]
        pk = self._pack_infer()
        with torch.cuda.stream(self.stream):
            if out is None:
                out = torch.empty(B, self.d_out, dtype=torch.float32, device=self.device)
            d = L.PdlpMlpInfer(rows=B, d_in=self.d_in, d_h=self.d_h, d_out=self.d_out, x=xb.data_ptr(),
                               w1p=pk["w1p"].data_ptr(), w2p=pk["w2p"].data_ptr(), w3p=pk["w3p"].data_ptr(),
                               b1=self._p("b1", self.w), b2=self._p("b2", self.w), b3=self._p("b3", self.w),
                               out=out.data_ptr())
            L.check(self.lib.pdlp_mlp_infer_bf16(C.byref(d), self.sh), "mlp infer")
        return out

    def loss_and_grad(self, xb: torch.Tensor, target: torch.Tensor, out_weight=None) -> torch.Tensor:
        d = self._desc(xb, target, out_weight)
        self._forward_desc(d)
        with torch.cuda.stream(self.stream):
            L.check(self.lib.pdlp_mlp_bwd_bf16(C.byref(d), self.sh), "mlp bwd")
        return self.loss

    def adam(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        self.step += 1
        with torch.cuda.stream(self.stream):
            L.check(self.lib.pdlp_adam_step(self.w.data_ptr(), self.g.data_ptr(), self.m.data_ptr(),
                                            self.v.data_ptr(), self.wb.data_ptr(), self.n_params, lr, beta1, beta2,
                                            eps, self.step, self.sh), "adam")
        self._packed = self._packed_dz = None

    def save(self, path: str, extra: Optional[dict] = None):
        """Weights (and extra arrays, e.g. target normalisation) as safetensors."""
        from safetensors.numpy import save_file
        t = {k: np.ascontiguousarray(v.astype(np.float32)) for k, v in self.params_numpy().items()}
        for k, v in (extra or {}).items():
            t[k] = np.ascontiguousarray(np.asarray(v, np.float32))
        save_file(t, path)

    @staticmethod
    def load(path: str):
        from safetensors.numpy import load_file
        t = load_file(path)
        m = WarmStartMLP(t["W1"].shape[0], t["W1"].shape[1], t["W3"].shape[1])
        m.set_params({k: t[k] for k in ("W1", "b1", "W2", "b2", "W3", "b3")})
        extra = {k: v for k, v in t.items() if k not in m.offsets}
        return m, extra
