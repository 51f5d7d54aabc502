"""Pin the FP64 MLP oracle by central finite differences (no GPU)."""

import numpy as np

from oracle import mlp_ref


def _params(rng, di, dh, do):
    return dict(W1=rng.standard_normal((di, dh)) * 0.5, b1=rng.standard_normal(dh) * 0.1,
                W2=rng.standard_normal((dh, dh)) * 0.3, b2=rng.standard_normal(dh) * 0.1,
                W3=rng.standard_normal((dh, do)) * 0.3, b3=rng.standard_normal(do) * 0.1)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    B, di, dh, do = 5, 4, 7, 3
    P = _params(rng, di, dh, do)
    x = rng.standard_normal((B, di))
    t = rng.standard_normal((B, do))
    w = rng.uniform(0.5, 2.0, do)
    loss, g, _ = mlp_ref.loss_and_grads(x, t, **P, w=w)
    h = 1e-6
    for name, val in P.items():
        flat = val.reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp = mlp_ref.loss_and_grads(x, t, **P, w=w)[0]
            flat[idx] = old - h
            lm = mlp_ref.loss_and_grads(x, t, **P, w=w)[0]
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[name].reshape(-1)[idx]) <= 1e-6 * (1 + abs(fd)), (name, idx, fd)


def test_bf16_round_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-9, -2.5, 1e-3], np.float32)
    r = mlp_ref.bf16_round(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0                       # halfway, rounds to even (mantissa 0)
    assert r[2] == 1.0 + 2**-7               # above halfway rounds up
    assert r[3] == -2.5
    assert abs(r[4] - 1e-3) <= 2**-9 * 1e-3
