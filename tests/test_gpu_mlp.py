"""GPU checks of the BF16-MFMA MLP kernels against the FP64 oracle (needs an MI355X)."""

import ctypes as C

import numpy as np
import pytest
import torch

from oracle import mlp_ref
from sovereign_solver_amd.engine import _lib as L
from sovereign_solver_amd.learned.mlp import WarmStartMLP

pytestmark = pytest.mark.gpu


def _bf(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev).to(torch.bfloat16).contiguous()


# (64, 64, 32), (256, 512, 224), (200, 136, 72): leading dimensions multiples of 8
# (16-byte loads; 200 and 136 leave partial tiles); the others 2-byte loads
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (70, 131, 45), (5, 300, 17), (256, 512, 224), (200, 136, 72)])
def test_gemm_matches_fp64(gpu, ta, tb, M, N, K):
    rng = np.random.default_rng(M * 7 + N + K)
    A = mlp_ref.bf16_round(rng.standard_normal((M, K)))
    B = mlp_ref.bf16_round(rng.standard_normal((K, N)))
    bias = rng.standard_normal(N).astype(np.float32)
    At = _bf(A.T if ta else A, gpu)
    Bt = _bf(B.T if tb else B, gpu)
    bt = torch.from_numpy(bias).to(gpu)
    Ct = torch.zeros(M, N, dtype=torch.float32, device=gpu)
    g = L.PdlpGemm(M=M, N=N, K=K, lda=(M if ta else K), ldb=(K if tb else N), ldc=N, trans_a=ta, trans_b=tb,
                   relu=0, accumulate=0, alpha=1.0, A=At.data_ptr(), B=Bt.data_ptr(), bias=bt.data_ptr(),
                   mask=0, C=Ct.data_ptr(), Cb=0)
    L.check(L.lib().pdlp_gemm_bf16(C.byref(g), None), "gemm")
    torch.cuda.synchronize()
    want = A @ B + bias
    got = Ct.cpu().numpy()
    bound = (K + 2) * 2.0**-23 * (np.abs(A) @ np.abs(B) + np.abs(bias)) + 1e-6
    assert np.all(np.abs(got - want) <= bound), np.max(np.abs(got - want) / bound)


@pytest.mark.parametrize("n", [64, 63, 136])
def test_gemm_identity_pins_lane_layout(gpu, n):
    """A x I = A and I x B = B: a transposed or shifted operand/result layout breaks
    these (n = 64, 136: 16-byte loads; 63: 2-byte loads)."""
    rng = np.random.default_rng(3)
    A = mlp_ref.bf16_round(rng.standard_normal((n, n)))
    I = np.eye(n)
    for left, right in ((A, I), (I, A)):
        At, Bt = _bf(left, gpu), _bf(right, gpu)
        Ct = torch.zeros(n, n, dtype=torch.float32, device=gpu)
        g = L.PdlpGemm(M=n, N=n, K=n, lda=n, ldb=n, ldc=n, alpha=1.0, A=At.data_ptr(), B=Bt.data_ptr(),
                       C=Ct.data_ptr())
        L.check(L.lib().pdlp_gemm_bf16(C.byref(g), None), "gemm")
        torch.cuda.synchronize()
        assert np.array_equal(Ct.cpu().numpy(), A)


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
def test_gemm_padded_leading_dims(gpu, ta, tb):
    """lda, ldb multiples of 8 with K = 45 and M, N = 70, 131 inside them: 16-byte
    loads with partial vectors at the edges."""
    M, N, K = 70, 131, 45
    rng = np.random.default_rng(11)
    A = mlp_ref.bf16_round(rng.standard_normal((M, K)))
    B = mlp_ref.bf16_round(rng.standard_normal((K, N)))
    As = np.zeros((K, 72) if ta else (M, 48))
    Bs = np.zeros((N, 48) if tb else (K, 136))
    if ta:
        As[:, :M] = A.T
    else:
        As[:, :K] = A
    if tb:
        Bs[:, :K] = B.T
    else:
        Bs[:, :N] = B
    At, Bt = _bf(As, gpu), _bf(Bs, gpu)
    Ct = torch.zeros(M, N, dtype=torch.float32, device=gpu)
    g = L.PdlpGemm(M=M, N=N, K=K, lda=As.shape[1], ldb=Bs.shape[1], ldc=N, trans_a=ta, trans_b=tb, alpha=1.0,
                   A=At.data_ptr(), B=Bt.data_ptr(), C=Ct.data_ptr())
    L.check(L.lib().pdlp_gemm_bf16(C.byref(g), None), "gemm")
    torch.cuda.synchronize()
    want = A @ B
    bound = (K + 2) * 2.0**-23 * (np.abs(A) @ np.abs(B)) + 1e-6
    assert np.all(np.abs(Ct.cpu().numpy() - want) <= bound)


@pytest.mark.parametrize("split,K", [(16, 4000), (27, 14025), (3, 130)])
def test_gemm_split_k(gpu, split, K):
    """Split K (dz2's shape: B^T operand, ReLU-derivative mask, BF16 output): within
    the FP32 bound of the FP64 product, and the same bits on a second run."""
    M, N = 256, 512
    rng = np.random.default_rng(K)
    A = mlp_ref.bf16_round(rng.standard_normal((M, K)))
    B = mlp_ref.bf16_round(rng.standard_normal((K, N)))
    mask = (rng.random((M, N)) > 0.3).astype(np.float64)
    At, Bt, Mt = _bf(A, gpu), _bf(B.T, gpu), _bf(mask, gpu)
    work = torch.zeros(split * M * N, dtype=torch.float32, device=gpu)
    outs = []
    for _ in range(2):
        Ct = torch.zeros(M, N, dtype=torch.float32, device=gpu)
        Cb = torch.zeros(M, N, dtype=torch.bfloat16, device=gpu)
        g = L.PdlpGemm(M=M, N=N, K=K, lda=K, ldb=K, ldc=N, trans_a=0, trans_b=1, alpha=1.0, split_k=split,
                       A=At.data_ptr(), B=Bt.data_ptr(), mask=Mt.data_ptr(), C=Ct.data_ptr(), Cb=Cb.data_ptr(),
                       work=work.data_ptr())
        L.check(L.lib().pdlp_gemm_bf16(C.byref(g), None), "gemm")
        torch.cuda.synchronize()
        outs.append((Ct.cpu().numpy(), Cb.float().cpu().numpy()))
    want = (A @ B) * mask
    bound = (K + 2) * 2.0**-23 * (np.abs(A) @ np.abs(B)) + 1e-6
    assert np.all(np.abs(outs[0][0] - want) <= bound)
    np.testing.assert_allclose(outs[0][1], outs[0][0], rtol=2.0**-8)      # BF16 copy of the same values
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    # split_k without a work space is refused
    g.work = 0
    with pytest.raises(L.PdlpError):
        L.check(L.lib().pdlp_gemm_bf16(C.byref(g), None), "gemm")


def test_mlp_forward_backward_vs_oracle(gpu):
    rng = np.random.default_rng(0)
    B, di, dh, do = 96, 24, 64, 200
    mlp = WarmStartMLP(di, dh, do, seed=1)
    P = mlp.params_numpy()
    P["b1"] = rng.standard_normal(dh) * 0.1
    P["b3"] = rng.standard_normal(do) * 0.1
    mlp.set_params(P)
    x = rng.standard_normal((B, di))
    t = rng.standard_normal((B, do))
    w = rng.uniform(0.5, 2.0, do)
    xb = _bf(mlp_ref.bf16_round(x), gpu)
    tt = torch.from_numpy(t.astype(np.float32)).to(gpu)
    wt = torch.from_numpy(w.astype(np.float32)).to(gpu)
    loss_t = mlp.loss_and_grad(xb, tt, wt)
    mlp.stream.synchronize()                 # the loss is written on the network's stream
    loss = float(loss_t.item())
    G = mlp.grads_numpy()
    P32 = {k: v.astype(np.float32).astype(np.float64) for k, v in mlp.params_numpy().items()}
    ref_loss, ref_g, ref_out = mlp_ref.loss_and_grads(x, t.astype(np.float32).astype(np.float64), **P32, w=w,
                                                      rnd=mlp_ref.bf16_round)
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    errs = {k: np.linalg.norm(G[k] - ref_g[k]) / max(np.linalg.norm(ref_g[k]), 1e-30) for k in ref_g}
    print("relative gradient errors:", {k: f"{v:.2e}" for k, v in errs.items()})
    # measured at most 8e-8: at these sizes no activation crosses a BF16 rounding boundary
    # between the FP32 and FP64 sums, so what is left is FP32 summation order
    for k, e in errs.items():
        assert e < 1e-6, (k, e)


def test_adam_step_matches_formula(gpu):
    mlp = WarmStartMLP(8, 16, 4, seed=2)
    w0 = mlp.w.cpu().numpy().astype(np.float64)
    g = np.random.default_rng(1).standard_normal(mlp.n_params).astype(np.float32)
    mlp.g.copy_(torch.from_numpy(g).to(gpu))
    mlp.adam(lr=1e-2)
    mlp.stream.synchronize()
    m = 0.1 * g
    v = 0.001 * g.astype(np.float64) ** 2
    want = w0 - 1e-2 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    np.testing.assert_allclose(mlp.w.cpu().numpy(), want, rtol=1e-5, atol=1e-6)
    wb = mlp.wb.float().cpu().numpy()
    np.testing.assert_allclose(wb, mlp_ref.bf16_round(mlp.w.cpu().numpy()), rtol=0, atol=0)


def test_training_reduces_loss(gpu):
    rng = np.random.default_rng(5)
    X = rng.standard_normal((2048, 16))
    Wt = rng.standard_normal((16, 40))
    T = np.tanh(X @ Wt) + 0.5 * (X[:, :1] ** 2)
    from sovereign_solver_amd.learned.warmstart import train
    _, _, _, losses = train(X, T, d_h=64, epochs=15, batch=256, lr=3e-3, step_weight=1.0)
    assert losses[-1] < 0.25 * losses[0], losses


def _integer_net(rng, d_in, d_out, d_h=512):
    """Small-integer inputs, weights and biases: every product and partial sum of the
    three layers is an integer below 2^24 (exact in FP32 in any order) and the hidden
    activations' BF16 rounding is the same for any summation order, so the fused
    kernel must give the oracle's values bit for bit."""
    P = {"W1": rng.integers(-1, 2, (d_in, d_h)) * (rng.random((d_in, d_h)) < 0.5),
         "b1": rng.integers(-3, 4, d_h),
         "W2": rng.integers(-1, 2, (d_h, d_h)) * (rng.random((d_h, d_h)) < 0.1),
         "b2": rng.integers(-3, 4, d_h),
         "W3": rng.integers(-1, 2, (d_h, d_out)) * (rng.random((d_h, d_out)) < 0.5),
         "b3": rng.integers(-3, 4, d_out)}
    return {k: v.astype(np.float64) for k, v in P.items()}


@pytest.mark.parametrize("rows,d_in,d_out", [(128, 48, 8), (300, 88, 8), (77, 8, 1), (1000, 128, 32), (129, 40, 5)])
def test_fused_inference_exact_on_integers(gpu, rows, d_in, d_out):
    """pdlp_mlp_infer_bf16 against the FP64 oracle with the kernels' BF16 rounding points,
    bit for bit on integer data: pins the packed operand orders (W1^T, W2^T, the layer-3
    gather), the permuted k slots of layer 2, the LDS row layout, the wavefront
    reduction and the row tail (rows not a multiple of 128)."""
    rng = np.random.default_rng(rows * 31 + d_in + d_out)
    mlp = WarmStartMLP(d_in, 512, d_out, seed=3)
    P = _integer_net(rng, d_in, d_out)
    mlp.set_params(P)
    x = rng.integers(-2, 3, (rows, d_in)).astype(np.float64)
    out = mlp.infer(_bf(x, gpu))
    mlp.stream.synchronize()
    _, _, want = mlp_ref.forward(x, **P, rnd=mlp_ref.bf16_round)
    assert np.array_equal(out.cpu().numpy().astype(np.float64), want)


def test_fused_inference_matches_layer_by_layer(gpu):
    """Random data: the fused kernel and the three-GEMM forward round the hidden
    activations at the same points; they differ only where an FP32 summation-order
    difference moves a value across a BF16 rounding boundary (one BF16 ulp, 2^-8
    relative, of that activation)."""
    rng = np.random.default_rng(11)
    d_in, d_out, rows = 88, 8, 4096
    mlp = WarmStartMLP(d_in, 512, d_out, seed=5)
    P = mlp.params_numpy()
    P["b1"] = rng.standard_normal(512) * 0.1
    P["b2"] = rng.standard_normal(512) * 0.1
    P["b3"] = rng.standard_normal(d_out) * 0.1
    mlp.set_params(P)
    x = rng.standard_normal((rows, d_in))
    xb = _bf(mlp_ref.bf16_round(x), gpu)
    fused_t = mlp.infer(xb)
    mlp.stream.synchronize()                 # outputs live on the network's stream
    fused = fused_t.cpu().numpy().astype(np.float64)
    layered_t = mlp.forward_layered(xb)
    mlp.stream.synchronize()
    layered = layered_t.cpu().numpy().astype(np.float64)
    P32 = {k: v.astype(np.float32).astype(np.float64) for k, v in mlp.params_numpy().items()}
    h1, h2, want = mlp_ref.forward(x, **P32, rnd=mlp_ref.bf16_round)
    W2b, W3b = mlp_ref.bf16_round(P32["W2"]), mlp_ref.bf16_round(P32["W3"])
    # every hidden activation off by at most one BF16 ulp, propagated (ReLU is 1-Lipschitz), plus FP32 sums
    dh1 = 2.0**-8 * np.abs(h1)
    dh2 = 2.0**-8 * np.abs(h2) + (dh1 @ np.abs(W2b)) * (1 + 2.0**-8)
    bound = dh2 @ np.abs(W3b) + 600 * 2.0**-24 * (np.abs(h2) @ np.abs(W3b) + np.abs(P32["b3"])) + 1e-6
    for got in (fused, layered):
        assert np.all(np.abs(got - want) <= bound), np.max(np.abs(got - want) / bound)
    # and in aggregate the two GPU paths agree far inside that bound
    assert np.linalg.norm(fused - layered) <= 1e-3 * np.linalg.norm(layered)


def test_fused_inference_refreshes_after_weight_update(gpu):
    rng = np.random.default_rng(2)
    mlp = WarmStartMLP(16, 512, 4, seed=1)
    P = _integer_net(rng, 16, 4)
    mlp.set_params(P)
    x = rng.integers(-2, 3, (64, 16)).astype(np.float64)
    xb = _bf(x, gpu)
    a_t = mlp.infer(xb)
    mlp.stream.synchronize()
    a = a_t.cpu().numpy()
    P2 = dict(P, W3=-P["W3"], b3=-P["b3"])
    mlp.set_params(P2)                       # the packed weights must follow
    b_t = mlp.infer(xb)
    mlp.stream.synchronize()
    b = b_t.cpu().numpy()
    assert np.array_equal(a, -b) and np.any(a != 0)


@pytest.mark.parametrize("rows,d_in", [(300, 88), (4096, 48)])
def test_fused_training_forward_stores_hidden_layers_exactly(gpu, rows, d_in):
    """The training forward of a d_h = 512 network runs the fused kernel and stores h1 and
    h2 for the backward pass: on integer data every stored activation and output equals
    the oracle's bit for bit (the row tail included)."""
    rng = np.random.default_rng(rows + d_in)
    mlp = WarmStartMLP(d_in, 512, 8, seed=4)
    assert mlp.can_infer_fused
    P = _integer_net(rng, d_in, 8)
    mlp.set_params(P)
    x = rng.integers(-2, 3, (rows, d_in)).astype(np.float64)
    out_t = mlp.forward(_bf(x, gpu))
    mlp.stream.synchronize()
    bu = mlp._buffers(rows)
    h1, h2, want = mlp_ref.forward(x, **P, rnd=mlp_ref.bf16_round)
    assert np.array_equal(out_t.cpu().numpy().astype(np.float64), want)
    assert np.array_equal(bu["h1"].float().cpu().numpy().astype(np.float64), h1)
    assert np.array_equal(bu["h2"].float().cpu().numpy().astype(np.float64), h2)


@pytest.mark.parametrize("fused", [True, False])
def test_node_shape_gradients_vs_oracle(gpu, fused):
    """A training step at the node model's shapes (d_in 88, d_h 512, d_out 8, 4,096 rows:
    fused forward, K-split weight gradients) against the FP64 oracle with the kernels'
    BF16 rounding points."""
    rng = np.random.default_rng(8)
    B, di, dh, do = 4096, 88, 512, 8
    mlp = WarmStartMLP(di, dh, do, seed=2)
    mlp.fused = fused                        # fused forward and backward kernels, or layer by layer
    P = mlp.params_numpy()
    P["b1"] = rng.standard_normal(dh) * 0.1
    P["b2"] = rng.standard_normal(dh) * 0.1
    P["b3"] = rng.standard_normal(do) * 0.1
    mlp.set_params(P)
    x = rng.standard_normal((B, di))
    t = rng.standard_normal((B, do))
    w = np.ones(do)
    xb = _bf(mlp_ref.bf16_round(x), gpu)
    tt = torch.from_numpy(t.astype(np.float32)).to(gpu)
    wt = torch.from_numpy(w.astype(np.float32)).to(gpu)
    loss_t = mlp.loss_and_grad(xb, tt, wt)
    mlp.stream.synchronize()                 # the loss is written on the network's stream
    loss = float(loss_t.item())
    G = mlp.grads_numpy()
    P32 = {k: v.astype(np.float32).astype(np.float64) for k, v in mlp.params_numpy().items()}
    ref_loss, ref_g, _ = mlp_ref.loss_and_grads(x, t.astype(np.float32).astype(np.float64), **P32, w=w,
                                                rnd=mlp_ref.bf16_round)
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    errs = {k: np.linalg.norm(G[k] - ref_g[k]) / max(np.linalg.norm(ref_g[k]), 1e-30) for k in ref_g}
    print("relative gradient errors:", {k: f"{v:.2e}" for k, v in errs.items()})
    # measured at most 8e-5 (W1): FP32 summation order and rare BF16 rounding flips of
    # the activations; a wrong operand, index or rounding point gives O(1)
    for k, e in errs.items():
        assert e < 1e-3, (k, e)
