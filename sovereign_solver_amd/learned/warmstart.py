"""Learned warm starts for a family of LPs that share one constraint matrix.

A *family* object provides ``problem(k) -> Problem`` (every instance holding
the same ``A_ub``/``A_eq`` objects), ``features_of(ids) -> [B, d_in]`` (the
network input) and the sizes ``n`` (columns) and ``m`` (rows ``[A_ub; A_eq]``).
:class:`sovereign_solver_amd.workloads.DispatchFamily` (varying demand and
cost) and :class:`sovereign_solver_amd.workloads.SetCoverNodes` (varying
bound fixings) are the two in this package.

Pipeline:

1. :func:`label` solves training instances with the shared-matrix batched
   engine from a cold start and keeps PDHG's own solutions (x*, y*) and a
   primal-weight label;
2. :class:`Normaliser` centres and scales inputs and targets;
3. :func:`train` fits :class:`WarmStartMLP` on the GPU (MFMA forward and
   backward kernels, Adam kernel);
4. :func:`solve_learned` runs the network on held-out instances and hands its
   output, still on the device, to the engine (``pdlp_family_set_start``:
   de-normalisation, projection onto the bounds and dual signs, tiling), then
   PDHG finishes and judges every solve, so a bad prediction only costs
   iterations.

Primal-weight label: with τ = ω τ̂ and σ = σ̂/ω, the PDHG distance bound
``‖z0 − z*‖²_M = ‖x0 − x*‖²_{T̂⁻¹}/ω + ω ‖y0 − y*‖²_{Σ̂⁻¹}`` is smallest at
``ω = ‖x0 − x*‖_{T̂⁻¹} / ‖y0 − y*‖_{Σ̂⁻¹}``; the label uses the cold-start
distances (x0 = y0 = 0) and the network predicts its logarithm.
"""

from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Optional

import numpy as np
import scipy.sparse as sp
import torch

from ..engine.family import FamilyBatch, FamilyEngine
from ..lp.batch import family_warm_start, results_of
from .mlp import WarmStartMLP


def step_label(K: sp.csr_matrix, x: np.ndarray, y: np.ndarray) -> float:
    """log ω* from a solution (x*, y*) (see the module docstring)."""
    aK = abs(K)
    colsum = np.maximum(np.asarray(aK.sum(axis=0)).ravel(), 1e-4)
    rowsum = np.maximum(np.asarray(aK.sum(axis=1)).ravel(), 1e-4)
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


def _engine_for(engines: dict, probs, tol: float, scheme: str) -> FamilyEngine:
    """A family engine for ``probs``: one built earlier for as many LPs is loaded with
    them (the matrix and the device buffers stay), else a new one."""
    eng = engines.get(len(probs))
    if eng is not None:
        eng.load(probs)
        return eng
    eng = engines[len(probs)] = FamilyEngine(FamilyBatch(probs), scheme=scheme)
    if len(probs) >= 256:
        eng.autotune(tol)
    return eng


def solve_family(fam, ids, tol: float, max_iters: int, warm=None, chunk: int = 2048, scheme: str = "adaptive"):
    """Batched solves of family instances (shared-matrix engine); returns
    SolveResults and the engine seconds (launch to results on the host)."""
    res = []
    secs = 0.0
    engines = {}
    for s in range(0, len(ids), chunk):
        part = ids[s: s + chunk]
        eng = _engine_for(engines, [fam.problem(int(k)) for k in part], tol, scheme)
        X, Y, w = family_warm_start(eng.fb, None if warm is None else warm[s: s + chunk])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out, xs, ys = eng.solve(tol, max_iters, X, Y, w)
        secs += time.perf_counter() - t0
        res += results_of(eng, out, xs, ys)
    return res, secs


def label(fam, ids, tol: float = 1e-4, max_iters: int = 100_000, chunk: int = 2048) -> Labels:
    ids = np.asarray(ids)
    res, secs = solve_family(fam, ids, tol, max_iters, chunk=chunk)
    X = np.stack([r.x for r in res]).astype(np.float32)
    Y = np.stack([r.y for r in res]).astype(np.float32)
    K = FamilyBatch([fam.problem(int(ids[0]))]).K_orig
    ls = np.array([step_label(K, r.x, r.y) for r in res])
    return Labels(ids, X, Y, ls, np.array([r.iterations for r in res]), np.array([r.status for r in res]), secs)


class Normaliser:
    """Per-dimension mean; scale per dimension (default) or one scale per block.

    ``blocks`` (list of slices) gives every dimension of a block the same scale,
    the RMS deviation of the whole block: a near-constant output (a variable
    at a bound, up to solver noise) is then not inflated to unit variance."""

    def __init__(self, a: np.ndarray, floor: float = 1e-6, blocks=None, identity: bool = False):
        if identity:                         # inputs already on a unit scale (e.g. +-1 fixings)
            self.mean, self.std = np.zeros(a.shape[1]), np.ones(a.shape[1])
            return
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


def output_blocks(fam):
    """x, y and step blocks of the network output."""
    return [slice(0, fam.n), slice(fam.n, fam.n + fam.m), slice(fam.n + fam.m, fam.n + fam.m + 1)]


def padded_width(d: int) -> int:
    """Output columns of the network for d targets: d rounded up to a multiple of 8."""
    return -(-d // 8) * 8


def train(feats: np.ndarray, targets: np.ndarray, d_h: int = 512, epochs: int = 20, batch: int = 512,
          lr: float = 1e-3, seed: int = 0, step_weight: float = 200.0, log=None, out_blocks=None,
          normalise_inputs: bool = True):
    """Fit the MLP on normalised (features, targets); returns (mlp, in_norm, out_norm, losses)."""
    dev = torch.device("cuda", torch.cuda.current_device())
    nin = Normaliser(feats, identity=not normalise_inputs)
    nout = Normaliser(targets, blocks=out_blocks) if out_blocks is not None else Normaliser(targets)
    X = torch.from_numpy(nin.fwd(feats).astype(np.float32)).to(dev).to(torch.bfloat16)
    # the output layer is padded to a multiple of 8 columns (zero targets, zero
    # loss weight): rows of 16 bytes keep every GEMM operand on 16-byte loads
    d_t = targets.shape[1]
    d_net = padded_width(d_t)
    T = torch.from_numpy(np.pad(nout.fwd(targets).astype(np.float32), ((0, 0), (0, d_net - d_t)))).to(dev)
    w = np.zeros(d_net, np.float32)
    w[:d_t] = 1.0
    w[d_t - 1] = step_weight                 # one scalar among many outputs
    W = torch.from_numpy(w).to(dev)
    mlp = WarmStartMLP(feats.shape[1], d_h, d_net, seed=seed, device=dev)
    rng = np.random.default_rng(seed)
    N = feats.shape[0]
    losses = []
    for ep in range(epochs):
        perm = rng.permutation(N)
        tot, nb = 0.0, 0
        lr_ep = lr * 0.5 * (1 + np.cos(np.pi * ep / max(1, epochs)))
        for s in range(0, N - batch + 1, batch):
            with torch.cuda.stream(mlp.stream):        # allocated on the stream that uses it
                idx = torch.from_numpy(perm[s: s + batch]).to(dev)
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


def predict(mlp: WarmStartMLP, nin: Normaliser, nout: Normaliser, feats: np.ndarray) -> np.ndarray:
    """De-normalised network outputs [B, n + m + 1] on the host."""
    with torch.cuda.stream(mlp.stream):
        X = torch.from_numpy(nin.fwd(feats).astype(np.float32)).to(mlp.device).to(torch.bfloat16).contiguous()
    out = mlp.forward(X)
    mlp.stream.synchronize()
    return nout.inv(out.cpu().numpy()[:, : nout.mean.size].astype(np.float64))


def warm_starts(pred: np.ndarray, n: int, m: int, use_x: bool = True, use_y: bool = True,
                use_step: bool = True) -> list:
    """(x0, y0, step) per instance from de-normalised outputs (the engine projects
    x onto the bounds and y onto its sign constraints)."""
    return [(pred[k, :n] if use_x else None, pred[k, n: n + m] if use_y else None,
             float(np.exp(np.clip(pred[k, -1], -8, 8))) if use_step else None) for k in range(len(pred))]


class LearnedStart:
    """A trained network with its normalisers, ready to start family solves on
    the device."""

    def __init__(self, mlp: WarmStartMLP, nin: Normaliser, nout: Normaliser):
        self.mlp, self.nin, self.nout = mlp, nin, nout
        dv = mlp.device
        self.in_mean = torch.from_numpy(nin.mean.astype(np.float32)).to(dv)
        self.in_inv_std = torch.from_numpy((1.0 / nin.std).astype(np.float32)).to(dv)
        pad = mlp.d_out - nout.mean.size        # padded output columns (unused by the engine)
        self.out_scale = torch.from_numpy(np.pad(nout.std, (0, pad), constant_values=1.0).astype(np.float32)).to(dv)
        self.out_shift = torch.from_numpy(np.pad(nout.mean, (0, pad)).astype(np.float32)).to(dv)

    def forward(self, feats: np.ndarray):
        """Network output [B, d_out] (normalised units) on the device, and an event
        marking its completion on the network's stream."""
        mlp = self.mlp
        with torch.cuda.stream(mlp.stream):
            f = torch.from_numpy(np.ascontiguousarray(feats, np.float32)).to(mlp.device, non_blocking=True)
        return self.forward_device(f)

    def forward_device(self, f: torch.Tensor):
        """As :meth:`forward` from features already in HBM ([B, d_in] FP32, allocated on
        or synchronised with the network's stream)."""
        mlp = self.mlp
        with torch.cuda.stream(mlp.stream):
            xb = ((f - self.in_mean) * self.in_inv_std).to(torch.bfloat16).contiguous()
        out = mlp.forward(xb)
        ev = torch.cuda.Event()
        ev.record(mlp.stream)
        return out, ev


def solve_learned(fam, ids, ls: LearnedStart, tol: float, max_iters: int, use_x: bool = True, use_y: bool = True,
                  use_step: bool = True, scheme: str = "adaptive", chunk: int = 2048):
    """Learned-warm-started batched solves; returns SolveResults and the seconds
    from the features on the host to the results on the host (network included)."""
    res = []
    secs = 0.0
    n, m = fam.n, fam.m
    engines = {}
    for s in range(0, len(ids), chunk):
        part = ids[s: s + chunk]
        eng = _engine_for(engines, [fam.problem(int(k)) for k in part], tol, scheme)
        feats = fam.features_of(part)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out, ev = ls.forward(feats)
        eng.start_from_device(tol, out, x_off=0 if use_x else -1, y_off=n if use_y else -1,
                              w_col=n + m if use_step else -1, scale=ls.out_scale, shift=ls.out_shift, ready=ev)
        eng.run(max_iters)
        o, xs, ys = eng.results()
        secs += time.perf_counter() - t0
        res += results_of(eng, o, xs, ys)
    return res, secs
