"""GPU checks of the BF16 MFMA MLP (nrpyell_mlp_*) against the FP64 restatement in
oracle/mlp_ref.py, which rounds activations and back-propagated gradients to BF16 where the
kernels store them; the tolerances cover FP32-vs-FP64 accumulation and occasional BF16
rounding flips (2^-8 relative) of individual activations."""
import numpy as np
import pytest

from nrpyelliptic_amd import mlp as nmlp
from nrpyelliptic_amd.runtime import DeviceArray
from oracle import mlp_ref

pytestmark = pytest.mark.gpu


def _inputs(M, seed):
    rng = np.random.default_rng(seed)
    X = np.zeros((M, nmlp.K0), np.float32)
    X[:, :23] = rng.uniform(-1, 1, (M, 23))
    bits = nmlp.bf16_bits(X)
    return bits, nmlp.bf16_to_float(bits).astype(np.float64), rng.uniform(-1, 1, M)


def test_forward_matches_fp64_reference():
    net = nmlp.InitialGuessMLP(seed=11)
    p = net.get_params().astype(np.float64)
    bits, X, _ = _inputs(256, 1)
    x = DeviceArray(bits.size, np.uint16).upload(bits.ravel())
    y = net.forward(x, 256)
    y_ref, _, _ = mlp_ref.forward(p, X)
    assert np.max(np.abs(y - y_ref)) <= 5e-3 * np.max(np.abs(y_ref)), np.max(np.abs(y - y_ref))


# the weight-gradient reduction runs over 32 row ranges (multiples of 16 rows): 512 rows give 32
# ranges of 16; 64 rows four ranges of 16 and 28 empty ones; 576 rows 18 ranges of 32 and 14 empty
@pytest.mark.parametrize("M", [512, 64, 576])
def test_gradients_match_fp64_reference(M):
    net = nmlp.InitialGuessMLP(seed=12)
    p = net.get_params().astype(np.float64)
    bits, X, t = _inputs(M, 2)
    x = DeviceArray(bits.size, np.uint16).upload(bits.ravel())
    td = DeviceArray(t.size, np.float32).upload(t.astype(np.float32))
    loss, g = net.loss_grad(x, td, M)
    loss_ref, g_ref = mlp_ref.loss_grad(p, X, t.astype(np.float32).astype(np.float64))
    assert loss == pytest.approx(loss_ref, rel=1e-3)
    # per-tensor relative error of the gradient
    off = 0
    for l in range(1, 6):
        k = 64 if l == 1 else 256
        for size in (256 * k, 256):
            a, b = g[off:off + size], g_ref[off:off + size]
            assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-12, (l, size)
            off += size
    a, b = g[off:], g_ref[off:]
    assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b)


def test_gradients_are_deterministic():
    """The weight-gradient partials are summed in a fixed order: repeated calls agree bit for bit."""
    net = nmlp.InitialGuessMLP(seed=14)
    bits, _, t = _inputs(1024, 5)
    x = DeviceArray(bits.size, np.uint16).upload(bits.ravel())
    td = DeviceArray(t.size, np.float32).upload(t.astype(np.float32))
    _, g1 = net.loss_grad(x, td, 1024)
    _, g2 = net.loss_grad(x, td, 1024)
    assert np.array_equal(g1, g2)


def test_adam_step_applies_the_last_gradient():
    """nrpyell_mlp_adam_step after loss_grad: the first Adam step moves every parameter with a
    nonzero gradient by lr * g / (|g| + eps) ~ lr sign(g)."""
    net = nmlp.InitialGuessMLP(seed=15)
    p0 = net.get_params()
    bits, _, t = _inputs(256, 6)
    x = DeviceArray(bits.size, np.uint16).upload(bits.ravel())
    td = DeviceArray(t.size, np.float32).upload(t.astype(np.float32))
    _, g = net.loss_grad(x, td, 256)
    check_lr = 1e-3
    from nrpyelliptic_amd.runtime import check
    check(nmlp._bind().nrpyell_mlp_adam_step(net.h, check_lr), "adam_step")
    dp = net.get_params() - p0
    big = np.abs(g) > 1e-5
    np.testing.assert_allclose(dp[big], -check_lr * np.sign(g[big]), rtol=2e-2, atol=0)


def test_training_reduces_loss():
    net = nmlp.InitialGuessMLP(seed=13)
    rng = np.random.default_rng(3)
    N = 1 << 15
    X = np.zeros((N, nmlp.K0), np.float32)
    X[:, :4] = rng.uniform(-1, 1, (N, 4))
    t = (np.sin(2 * X[:, 0]) * np.cos(X[:, 1]) + 0.5 * X[:, 2] * X[:, 3]).astype(np.float32)
    x = DeviceArray(X.size, np.uint16).upload(nmlp.bf16_bits(X).ravel())
    td = DeviceArray(N, np.float32).upload(t)
    losses = net.train(x, td, N, batch=4096, steps=400, lr=2e-3, seed=5)
    assert np.all(np.isfinite(losses))
    assert losses[-20:].mean() < 0.02 * losses[:5].mean(), (losses[:5].mean(), losses[-20:].mean())
