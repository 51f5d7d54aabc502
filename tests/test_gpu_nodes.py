"""GPU: node-model kernels and the node-model warm start (needs an MI355X).

* ``pdlp_node_gather`` against the NumPy restatement of tests/test_nodes_host.py
  (BF16 outputs bit-exact against round-to-nearest-even, FP32 exact), with the
  shared and second sources, an explicit row list and hashed sampling;
* ``pdlp_node_assemble`` against the NumPy restatement;
* the backward pass with the K-split work space (a batch of 4,096 rows and
  8 outputs) against the FP64 oracle and against the unsplit pass;
* a two-stage node model on a small dispatch family: every warm-started solve
  stops at the tolerance by the host yardstick, with fewer iterations than cold.
"""

import ctypes as C

import numpy as np
import pytest
import torch

from oracle import mlp_ref, yardstick
from sovereign_solver_amd.engine import _lib as L
from sovereign_solver_amd.learned import nodes as N
from sovereign_solver_amd.learned import warmstart as W
from sovereign_solver_amd.learned.mlp import WarmStartMLP
from sovereign_solver_amd.workloads import DispatchFamily
from tests.parity import assert_objective_parity
from tests.test_nodes_host import assemble_ref, gather_ref

pytestmark = pytest.mark.gpu


def _t(a, dev, dt=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(dt)


def _gather(dev, src, shared, table, out, scale=None, mean=None, inv_std=None, src2=None, rows=None, count=0,
            seed=0):
    keep = [src, shared, table, scale, mean, inv_std, src2, rows]
    p = lambda t: 0 if t is None else t.data_ptr()
    g = L.PdlpNodeGather(batch=src.shape[0], nodes=table.shape[0], width=table.shape[1], ld_out=out.shape[1],
                         src_len=src.shape[1], shared_len=0 if shared is None else shared.numel(),
                         src2_len=0 if src2 is None else src2.shape[1], count=count, seed=seed, src=p(src),
                         shared=p(shared), src2=p(src2), table=p(table), scale=p(scale), mean=p(mean),
                         inv_std=p(inv_std), rows=p(rows))
    if out.dtype == torch.bfloat16:
        g.out_bf16 = out.data_ptr()
    else:
        g.out_f32 = out.data_ptr()
    L.check(L.lib().pdlp_node_gather(C.byref(g), None), "gather")
    torch.cuda.synchronize()
    del keep


@pytest.mark.parametrize("ld_out", [8, 48, 72])
def test_gather_matches_reference(gpu, ld_out):
    rng = np.random.default_rng(ld_out)
    B, nodes, Ls, Lsh, L2 = 5, 37, 23, 6, 11
    width = min(ld_out - 1, 45)
    src = rng.standard_normal((B, Ls)).astype(np.float32)
    shared = rng.standard_normal(Lsh).astype(np.float32)
    src2 = rng.standard_normal((B, L2)).astype(np.float32)
    table = rng.integers(-1, Ls + Lsh + L2, size=(nodes, width)).astype(np.int32)
    scale = rng.uniform(0.5, 2.0, (nodes, width)).astype(np.float32)
    mean = rng.standard_normal(width).astype(np.float32)
    inv = rng.uniform(0.5, 2.0, width).astype(np.float32)
    want = gather_ref(src, shared, table, scale, mean, inv, src2=src2, ld_out=ld_out)
    T = dict(src=_t(src, gpu), shared=_t(shared, gpu), table=_t(table, gpu, torch.int32), scale=_t(scale, gpu),
             mean=_t(mean, gpu), inv_std=_t(inv, gpu), src2=_t(src2, gpu))
    out = torch.full((B * nodes, ld_out), 7.0, device=gpu)
    _gather(gpu, out=out, **T)
    # FP32 arithmetic in the kernel, FP64 in the reference
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=2e-6, atol=2e-6)
    ob = torch.empty((B * nodes, ld_out), dtype=torch.bfloat16, device=gpu)
    _gather(gpu, out=ob, **T)
    np.testing.assert_array_equal(ob.float().cpu().numpy(), out.to(torch.bfloat16).float().cpu().numpy())
    rows = rng.integers(0, B * nodes, 50)
    orow = torch.zeros((50, ld_out), device=gpu)
    _gather(gpu, out=orow, rows=_t(rows, gpu, torch.int64), count=50, **T)
    np.testing.assert_allclose(orow.cpu().numpy(), want[rows], rtol=2e-6, atol=2e-6)


def test_gather_hashed_rows_pair_inputs_and_targets(gpu):
    # column 0 reads src[b][0] = b, column 1 reads shared[j] = j: each row names its (b, j)
    B, nodes = 9, 300
    src = np.zeros((B, 1), np.float32)
    src[:, 0] = np.arange(B)
    shared = np.arange(nodes, dtype=np.float32)
    table = np.stack([np.zeros(nodes), 1 + np.arange(nodes)], axis=1).astype(np.int32)
    T = dict(src=_t(src, gpu), shared=_t(shared, gpu), table=_t(table, gpu, torch.int32))
    R = 20000
    a = torch.zeros((R, 8), device=gpu)
    b = torch.zeros((R, 8), device=gpu)
    c = torch.zeros((R, 8), device=gpu)
    _gather(gpu, out=a, count=R, seed=12345, **T)
    _gather(gpu, out=b, count=R, seed=12345, **T)
    _gather(gpu, out=c, count=R, seed=999, **T)
    A, Bv, Cv = (t.cpu().numpy() for t in (a, b, c))
    assert np.array_equal(A, Bv)                             # the same seed draws the same rows
    assert np.mean(np.all(A[:, :2] == Cv[:, :2], axis=1)) < 0.01
    q = A[:, 0].astype(int) * nodes + A[:, 1].astype(int)
    assert q.min() >= 0 and q.max() < B * nodes
    # uniform over the 2,700 rows: every instance drawn about R / B times
    counts = np.bincount(A[:, 0].astype(int), minlength=B)
    assert np.all(np.abs(counts - R / B) < 6 * np.sqrt(R / B))


def test_assemble_matches_reference(gpu):
    rng = np.random.default_rng(2)
    B, nodes, ld_net, length = 6, 40, 8, 55
    cnt = rng.integers(0, 4, length)
    indptr = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    nnz = int(indptr[-1])
    node = rng.integers(0, nodes, nnz).astype(np.int32)
    chan = rng.integers(0, 4, nnz).astype(np.int32)
    weight = rng.standard_normal(nnz).astype(np.float32)
    bias = rng.standard_normal(length).astype(np.float32)
    net = rng.standard_normal((B * nodes, ld_net)).astype(np.float32)
    want = assemble_ref(net.astype(np.float64), nodes, indptr, node, chan, weight, bias, length)
    out = torch.full((B, length + 3), -5.0, device=gpu)
    T = [_t(indptr, gpu, torch.int32), _t(node, gpu, torch.int32), _t(chan, gpu, torch.int32), _t(weight, gpu),
         _t(bias, gpu), _t(net, gpu)]
    a = L.PdlpNodeAssembly(batch=B, nodes=nodes, ld_net=ld_net, len=length, ld_out=length + 3, nnz=nnz,
                           indptr=T[0].data_ptr(), node=T[1].data_ptr(), chan=T[2].data_ptr(), weight=T[3].data_ptr(),
                           bias=T[4].data_ptr(), net=T[5].data_ptr(), out=out.data_ptr())
    L.check(L.lib().pdlp_node_assemble(C.byref(a), None), "assemble")
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    np.testing.assert_allclose(got[:, :length], want, rtol=1e-5, atol=1e-5)
    assert np.all(got[:, length:] == -5.0)                   # columns past len untouched


def test_backward_with_work_space_matches_oracle_and_unsplit(gpu):
    rng = np.random.default_rng(4)
    B, di, dh, do = 4096, 48, 256, 8
    mlp = WarmStartMLP(di, dh, do, seed=3)
    x = rng.standard_normal((B, di))
    t = rng.standard_normal((B, do))
    xb = torch.from_numpy(mlp_ref.bf16_round(x).astype(np.float32)).to(gpu).to(torch.bfloat16).contiguous()
    tt = torch.from_numpy(t.astype(np.float32)).to(gpu)
    assert B >= mlp.SPLIT_BATCH
    loss = float(mlp.loss_and_grad(xb, tt).item())
    G_split = mlp.grads_numpy()
    assert mlp._bufs[B]["work_floats"] > 0
    P32 = {k: v.astype(np.float32).astype(np.float64) for k, v in mlp.params_numpy().items()}
    ref_loss, ref_g, _ = mlp_ref.loss_and_grads(x, t.astype(np.float32).astype(np.float64), **P32,
                                                rnd=mlp_ref.bf16_round)
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    for k in ref_g:
        err = np.linalg.norm(G_split[k] - ref_g[k]) / max(np.linalg.norm(ref_g[k]), 1e-30)
        assert err < 2e-2, (k, err)
    # the same pass without the work space (one K part per weight gradient)
    mlp._bufs[B]["work_floats"] = 0
    mlp.loss_and_grad(xb, tt)
    G_one = mlp.grads_numpy()
    for k in G_one:
        np.testing.assert_allclose(G_split[k], G_one[k], rtol=1e-3, atol=1e-6 * np.abs(G_one[k]).max())
    # and the split pass gives the same bits twice (fixed summation order)
    mlp._bufs[B]["work_floats"] = 32 * dh * dh
    mlp.loss_and_grad(xb, tt)
    G2 = mlp.grads_numpy()
    assert all(np.array_equal(G2[k], G_split[k]) for k in G2)


def test_two_stage_node_model_on_a_small_dispatch_family(gpu):
    fam = DispatchFamily(G=40, T=12)
    train_ids, eval_ids = np.arange(2048), 10**6 + np.arange(96)
    lab = W.label(fam, train_ids)
    assert np.all(lab.status == "optimal")
    model = N.train_node_model(fam, lab, stages=2, steps=2500, rows=16384, calib_ids=train_ids[-64:])
    s1, s2 = N.stage_list(model)
    assert s2.prev is s1 and s2.width == s1.width + 5 * 4
    cold, _ = W.solve_family(fam, eval_ids, 1e-4, 100_000)
    warm, _ = N.solve_nodes(fam, eval_ids, model, 1e-4, 100_000)
    probs = [fam.problem(int(k)) for k in eval_ids]
    for p, r in zip(probs, warm):
        assert r.status == "optimal"
        assert yardstick.rel_kkt(p, r.x, r.y)[0] <= 1.02e-4
    ci = np.mean([r.iterations for r in cold])
    wi = np.mean([r.iterations for r in warm])
    print(f"small dispatch 40x12: cold mean {ci:.0f}, two-stage node model mean {wi:.0f}")
    assert wi < ci
    for k in range(4):
        assert_objective_parity(probs[k], warm[k].objective, x=warm[k].x, y=warm[k].y)
    # the starts are deterministic: the same network gives the same start twice
    V = fam.node_inputs(eval_ids[:8])
    a, _ = model.starts(V)
    b, _ = model.starts(V)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
