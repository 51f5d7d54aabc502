"""Learned warm starts for a parametric LP family (economic dispatch).

Pipeline:

1. :func:`label` solves training instances with the batched engine (cold
   start) and keeps PDHG's own solutions (x*, y*) and a step label;
2. :class:`Normaliser` z-scores inputs/targets per dimension;
3. :func:`train` fits :class:`WarmStartMLP` on the GPU (MFMA forward and
   backward kernels, Adam kernel);
4. :func:`predict_warm_start` turns network outputs into per-instance
   ``(x0, y0, step)`` for :func:`solve_lp_pdhg_batch`, projected onto the
   bounds and dual sign constraints;
5. :func:`evaluate` solves held-out instances cold and warm with the same
   engine, tolerance and iteration cap, so any gain is measured, and PDHG
   always finishes (and judges) every solve.

Step label: with τ = ω τ̂ and σ = σ̂/ω, the PDHG distance bound
``‖z0 − z*‖²_M = ‖x0 − x*‖²/ω_{T̂} + ω ‖y0 − y*‖²_{Σ̂}`` is smallest at
``ω = ‖x0 − x*‖_{T̂⁻¹} / ‖y0 − y*‖_{Σ̂⁻¹}``; the label uses the cold-start
distances (x0 = y0 = 0), and the network predicts its logarithm.
"""

from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np
import torch

from ..engine.batch import BatchEngine, PackedBatch
from ..lp.batch import pack_warm_start, results_of
from ..workloads import DispatchFamily
from .mlp import WarmStartMLP


def features_of(fam: DispatchFamily, ids) -> np.ndarray:
    return np.stack([fam.features(*fam.params(int(k))) for k in ids])


def step_label(fam: DispatchFamily, x: np.ndarray, y: np.ndarray) -> float:
    """log ω* from a solution (x*, y*) (see module docstring)."""
    A = abs(fam.A)
    colsum = np.maximum(np.asarray(A.sum(axis=0)).ravel(), 1e-4)
    rowsum = np.maximum(np.asarray(A.sum(axis=1)).ravel(), 1e-4)
    nx = np.sqrt(np.sum(x * x * colsum) / 0.999)
    ny = np.sqrt(np.sum(y * y * rowsum) / 0.999)
    return float(np.log(max(nx, 1e-12) / max(ny, 1e-12)))


@dataclass
class Labels:
    ids: np.ndarray
    x: np.ndarray          # [N, n]
    y: np.ndarray          # [N, m]
    log_step: np.ndarray   # [N]
    iterations: np.ndarray
    status: np.ndarray
    seconds: float


def solve_family(fam: DispatchFamily, ids, tol: float, max_iters: int, warm=None, chunk: int = 2048,
                 block: int = 0, scheme: str = "adaptive"):
    """Batched solves of family instances; returns SolveResults and engine seconds (launch to results)."""
    res = []
    secs = 0.0
    for s in range(0, len(ids), chunk):
        part = ids[s: s + chunk]
        probs = [fam.problem(int(k)) for k in part]
        pk = PackedBatch(probs)
        eng = BatchEngine(pk, block=block, scheme=scheme)
        x0, y0, w = pack_warm_start(pk, None if warm is None else warm[s: s + chunk])
        eng.set_warm_start(x0, y0, w)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.launch(tol, max_iters)
        out, xs, ys = eng.results()
        secs += time.perf_counter() - t0
        res += results_of(eng, out, xs, ys)
        del eng
    return res, secs


def label(fam: DispatchFamily, ids, tol: float = 1e-4, max_iters: int = 100_000) -> Labels:
    ids = np.asarray(ids)
    res, secs = solve_family(fam, ids, tol, max_iters)
    X = np.stack([r.x for r in res]).astype(np.float32)
    Y = np.stack([r.y for r in res]).astype(np.float32)
    ls = np.array([step_label(fam, r.x, r.y) for r in res])
    return Labels(ids, X, Y, ls, np.array([r.iterations for r in res]), np.array([r.status for r in res]), secs)


class Normaliser:
    """Per-dimension mean; scale per dimension (default) or one scale per block.

    ``blocks`` (list of slices) gives every dimension of a block the same scale,
    the RMS deviation of the whole block: a near-constant output (a variable
    at a bound, up to solver noise) is then not inflated to unit variance."""

    def __init__(self, a: np.ndarray, floor: float = 1e-6, blocks=None):
        self.mean = a.mean(axis=0)
        if blocks is None:
            self.std = np.maximum(a.std(axis=0), floor)
        else:
            self.std = np.empty(a.shape[1])
            for sl in blocks:
                d = a[:, sl] - self.mean[sl]
                self.std[sl] = max(float(np.sqrt(np.mean(d * d))), floor)

    def fwd(self, a):
        return (a - self.mean) / self.std

    def inv(self, a):
        return a * self.std + self.mean


def targets_of(lab: Labels) -> np.ndarray:
    return np.concatenate([lab.x, lab.y, lab.log_step[:, None]], axis=1)


def output_blocks(fam: DispatchFamily):
    """x, y and step blocks of the network output."""
    return [slice(0, fam.n), slice(fam.n, fam.n + fam.m), slice(fam.n + fam.m, fam.n + fam.m + 1)]


def train(feats: np.ndarray, targets: np.ndarray, d_h: int = 512, epochs: int = 20, batch: int = 512,
          lr: float = 1e-3, seed: int = 0, step_weight: float = 200.0, log=None, out_blocks=None):
    """Fit the MLP on normalised (features, targets); returns (mlp, in_norm, out_norm, losses)."""
    dev = torch.device("cuda", torch.cuda.current_device())
    nin = Normaliser(feats)
    nout = Normaliser(targets, blocks=out_blocks) if out_blocks is not None else Normaliser(targets)
    X = torch.from_numpy(nin.fwd(feats).astype(np.float32)).to(dev).to(torch.bfloat16)
    T = torch.from_numpy(nout.fwd(targets).astype(np.float32)).to(dev)
    w = np.ones(targets.shape[1], np.float32)
    w[-1] = step_weight                      # one scalar among ~14k outputs
    W = torch.from_numpy(w).to(dev)
    mlp = WarmStartMLP(feats.shape[1], d_h, targets.shape[1], seed=seed, device=dev)
    rng = np.random.default_rng(seed)
    N = feats.shape[0]
    losses = []
    for ep in range(epochs):
        perm = rng.permutation(N)
        tot, nb = 0.0, 0
        lr_ep = lr * 0.5 * (1 + np.cos(np.pi * ep / max(1, epochs)))
        for s in range(0, N - batch + 1, batch):
            idx = torch.from_numpy(perm[s: s + batch]).to(dev)
            with torch.cuda.stream(mlp.stream):
                xb = X.index_select(0, idx).contiguous()
                tb = T.index_select(0, idx).contiguous()
            loss = mlp.loss_and_grad(xb, tb, W)
            mlp.adam(lr=lr_ep)
            if nb % 20 == 0:
                mlp.stream.synchronize()
                tot += float(loss.item())
            nb += 1
        mlp.stream.synchronize()
        losses.append(tot / max(1, (nb + 19) // 20))
        if log:
            log(f"epoch {ep}: loss {losses[-1]:.4f}")
    return mlp, nin, nout, losses


def predict_warm_start(mlp: WarmStartMLP, nin: Normaliser, nout: Normaliser, fam: DispatchFamily, ids,
                       use_step: bool = True):
    feats = features_of(fam, ids)
    dev = mlp.device
    X = torch.from_numpy(nin.fwd(feats).astype(np.float32)).to(dev).to(torch.bfloat16).contiguous()
    out = mlp.forward(X)
    mlp.stream.synchronize()
    pred = nout.inv(out.cpu().numpy().astype(np.float64))
    n, m = fam.n, fam.m
    warm = []
    ub = np.tile(fam.pmax, fam.T)
    for k in range(len(ids)):
        x0 = np.clip(pred[k, :n], 0.0, ub)
        y0 = np.minimum(pred[k, n: n + m], 0.0)          # all rows are <=: y <= 0
        step = float(np.exp(np.clip(pred[k, -1], -8, 8))) if use_step else None
        warm.append((x0, y0, step))
    return warm


def evaluate(fam: DispatchFamily, ids, warm, tol: float = 1e-4, max_iters: int = 100_000,
             scheme: str = "adaptive"):
    """Cold vs warm batched solves of the same instances."""
    cold, t_cold = solve_family(fam, ids, tol, max_iters, scheme=scheme)
    hot, t_hot = solve_family(fam, ids, tol, max_iters, warm=warm, scheme=scheme)
    return dict(
        cold_iters=np.array([r.iterations for r in cold]), warm_iters=np.array([r.iterations for r in hot]),
        cold_ok=sum(r.status == "optimal" for r in cold), warm_ok=sum(r.status == "optimal" for r in hot),
        cold_s=t_cold, warm_s=t_hot, cold=cold, warm=hot)
