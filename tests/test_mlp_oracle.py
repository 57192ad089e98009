"""Pins the FP64 MLP restatement (oracle/mlp_ref.py) that checks the GPU MLP: its analytic
gradient equals central finite differences of its own loss (plain FP64 mode), and the BF16
rounding helper is round-to-nearest-even. CPU only."""
import numpy as np

from oracle import mlp_ref


def random_params(rng, scale=0.2):
    n = 0
    for l in range(1, mlp_ref.HIDDEN + 1):
        k = mlp_ref.K0 if l == 1 else mlp_ref.W
        n += mlp_ref.W * k + mlp_ref.W
    n += mlp_ref.W + 1
    return rng.normal(0, scale, n) / np.sqrt(64)


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = random_params(rng)
    X = rng.uniform(-1, 1, (8, mlp_ref.K0))
    t = rng.uniform(-1, 1, 8)
    _, g = mlp_ref.loss_grad(p, X, t, bf16_round=False)
    idx = rng.choice(p.size, 40, replace=False)
    idx = np.concatenate([idx, [p.size - 1, p.size - 2, 0, 64 * 256 + 3]])  # b_out, w_out, W_1, b_1
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = mlp_ref.loss_grad(p + e, X, t, bf16_round=False)
        lm, _ = mlp_ref.loss_grad(p - e, X, t, bf16_round=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5, 3.0e-3])
    r = mlp_ref.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0 + 2 ** -7 or r[1] == 1.0  # exactly half: ties to even
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == 1.0
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8
