"""FP64 restatement of the learned-initial-guess MLP (forward, loss, backprop).

TEST INFRASTRUCTURE: the checker for nrpyell_mlp_* (no reference counterpart; this network
exists only in this build). Same parameter layout as the library:
  for l = 1..5: W_l [256 x K_l] (K_1 = 64, else 256), b_l [256]; then w_out [256], b_out [1].
With bf16=True the activations H_l and the back-propagated dZ_l are rounded to BF16 exactly
where the GPU kernels store them in BF16, and weights are used as their BF16 copies, so the
remaining differences are FP32-vs-FP64 accumulation. With bf16=False it is a plain FP64
network, pinned by central finite differences (tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import numpy as np

K0, W, HIDDEN = 64, 256, 5


def bf16(x):
    x = np.asarray(x, dtype=np.float64).astype(np.float32)
    b = x.view(np.uint32)
    r = ((b >> 16) & 1) + np.uint32(0x7FFF)
    return (((b + r) >> 16) << 16).astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    Ws, bs, off = [], [], 0
    for l in range(1, HIDDEN + 1):
        k = K0 if l == 1 else W
        Ws.append(p[off:off + W * k].reshape(W, k))
        off += W * k
        bs.append(p[off:off + W])
        off += W
    w_out = p[off:off + W]
    off += W
    b_out = p[off]
    return Ws, bs, w_out, b_out


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p, X, bf16_round=True):
    Ws, bs, w_out, b_out = unpack(p)
    rnd = bf16 if bf16_round else (lambda a: a)
    H = rnd(X)
    Hs, Zs = [H], []
    for l in range(HIDDEN):
        Z = H @ rnd(Ws[l]).T + bs[l]
        H = rnd(Z * sigmoid(Z))
        Zs.append(Z)
        Hs.append(H)
    y = H @ w_out + b_out
    return y, Hs, Zs


def loss_grad(p, X, t, bf16_round=True):
    """Mean squared error and its gradient in the flat parameter layout."""
    Ws, bs, w_out, b_out = unpack(p)
    rnd = bf16 if bf16_round else (lambda a: a)
    y, Hs, Zs = forward(p, X, bf16_round)
    M = X.shape[0]
    diff = y - t
    loss = float(np.mean(diff ** 2))
    dy = 2.0 * diff / M
    g_w_out = Hs[-1].T @ dy
    g_b_out = dy.sum()

    def dsilu(z):
        s = sigmoid(z)
        return s * (1.0 + z * (1.0 - s))

    dZ = rnd(np.outer(dy, w_out) * dsilu(Zs[-1]))
    gWs, gbs = [None] * HIDDEN, [None] * HIDDEN
    for l in range(HIDDEN - 1, -1, -1):
        gWs[l] = dZ.T @ Hs[l]
        gbs[l] = dZ.sum(axis=0)
        if l > 0:
            dZ = rnd((dZ @ rnd(Ws[l])) * dsilu(Zs[l - 1]))
    parts = []
    for l in range(HIDDEN):
        parts += [gWs[l].ravel(), gbs[l]]
    parts += [g_w_out, np.array([g_b_out])]
    return loss, np.concatenate(parts)
