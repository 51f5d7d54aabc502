"""FP64 NumPy forward/backward of the warm-start MLP.  TEST INFRASTRUCTURE.

Network: h1 = relu(x W1 + b1), h2 = relu(h1 W2 + b2), out = h2 W3 + b3;
loss = Σ_{b,j} w_j (out_bj − t_bj)² / (B·d_out).  Weights are [fan_in][fan_out].

``forward_bf16like`` rounds the inputs, weights and stored activations to
BF16 exactly where the GPU kernels do, so the GPU gradients can be compared
against FP64 arithmetic on the same rounded operands.

Pinned by central finite differences in ``tests/test_mlp_oracle.py``.
"""

from __future__ import annotations

import numpy as np


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round FP32-representable values to BF16 (nearest even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def forward(x, W1, b1, W2, b2, W3, b3, rnd=lambda a: a):
    z1 = rnd(x) @ rnd(W1) + b1
    h1 = rnd(np.maximum(z1, 0.0))
    z2 = h1 @ rnd(W2) + b2
    h2 = rnd(np.maximum(z2, 0.0))
    out = h2 @ rnd(W3) + b3
    return h1, h2, out


def loss_and_grads(x, t, W1, b1, W2, b2, W3, b3, w=None, rnd=lambda a: a):
    """Loss and gradients; ``rnd`` mimics where the GPU rounds to BF16."""
    B, d_out = t.shape
    w = np.ones(d_out) if w is None else np.asarray(w, float)
    h1, h2, out = forward(x, W1, b1, W2, b2, W3, b3, rnd)
    diff = out - t
    inv = 1.0 / (B * d_out)
    loss = float(np.sum(w * diff * diff) * inv)
    dout = 2.0 * w * diff * inv
    gb3 = dout.sum(axis=0)
    dout_b = rnd(dout)
    gW3 = h2.T @ dout_b
    dz2 = rnd((dout_b @ rnd(W3).T) * (h2 > 0))
    gb2 = dz2.sum(axis=0)
    gW2 = h1.T @ dz2
    dz1 = rnd((dz2 @ rnd(W2).T) * (h1 > 0))
    gb1 = dz1.sum(axis=0)
    gW1 = rnd(x).T @ dz1
    return loss, dict(W1=gW1, b1=gb1, W2=gW2, b2=gb2, W3=gW3, b3=gb3), out
