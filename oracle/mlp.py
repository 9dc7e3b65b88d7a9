"""CPU oracle of the learned surrogate's network: forward pass, loss and gradient in FP64.

Test infrastructure only: tests/ use it to check the GPU training and inference kernels; no
product code imports it.

The surrogate has no counterpart in the reference repository, so nothing there pins it; this
module is pinned instead by central finite differences of its own loss (tests/test_mlp.py, mode
"exact") and by the closed forms of SiLU and the logistic loss.

Network (parameter vector layout shared with the library, row-major [out][in] weights):
    W1 128x3, b1 128, W2 128x128, b2 128, W3 128x128, b3 128, W4 5x128, b4 5   (34181 floats)
    a1 = silu(W1 x + b1), a2 = silu(W2 a1 + b2), a3 = silu(W3 a2 + b3), y = W4 a3 + b4
Loss of a batch of B rows with targets t and masks m (5 per row):
    sum_b [ sum_{o<3} w_o m_o (y_o - t_o)^2 / B  +  0.1/(2B) sum_{o=3,4} m_o (softplus(y_o) - t_o y_o) ]
    w = (1, 8, 8)
Modes:
    "exact"  FP64 throughout.
    "bf16"   the GPU's rounding points: the operands of the three 128-wide products (activations
             and weights, forward; upstream gradients and weights, backward; upstream gradients
             and activations, weight gradients) are rounded to bf16 (nearest even), straight-through
             in the backward pass; layer 1 and all biases stay unrounded.
"""
from __future__ import annotations

import numpy as np

H, NIN, NOUT = 128, 3, 5
W1, B1 = 0, 384
W2, B2 = 512, 512 + 16384
W3, B3 = B2 + 128, B2 + 128 + 16384
W4, B4 = B3 + 128, B3 + 128 + 640
NPARAMS = B4 + 5
assert NPARAMS == 34181
REG_W = np.array([1.0, 8.0, 8.0])


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) and back, via the float32 bit pattern."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p: np.ndarray) -> dict:
    p = np.asarray(p, dtype=np.float64)
    return dict(W1=p[W1:B1].reshape(H, NIN), b1=p[B1:W2], W2=p[W2:B2].reshape(H, H), b2=p[B2:W3],
                W3=p[W3:B3].reshape(H, H), b3=p[B3:W4], W4=p[W4:B4].reshape(NOUT, H), b4=p[B4:])


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sigmoid(z)


def dsilu(z):
    s = sigmoid(z)
    return s * (1.0 + z * (1.0 - s))


def softplus(z):
    return np.where(z > 0, z + np.log1p(np.exp(-np.abs(z))), np.log1p(np.exp(-np.abs(z))))


def forward(p, x, mode="bf16"):
    """Outputs (B, 5) and the intermediates of the pass."""
    q = unpack(p)
    r = bf16 if mode == "bf16" else (lambda a: np.asarray(a, dtype=np.float64))
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ q["W1"].T + q["b1"]
    a1 = silu(z1)
    z2 = r(a1) @ r(q["W2"]).T + q["b2"]
    a2 = silu(z2)
    z3 = r(a2) @ r(q["W3"]).T + q["b3"]
    a3 = silu(z3)
    y = r(a3) @ r(q["W4"]).T + q["b4"]
    return y, dict(x=x, z1=z1, a1=a1, z2=z2, a2=a2, z3=z3, a3=a3)


def loss_grad(p, x, t, m, mode="bf16"):
    """Loss and gradient (NPARAMS) of a batch; same definition as the GPU training step."""
    q = unpack(p)
    r = bf16 if mode == "bf16" else (lambda a: np.asarray(a, dtype=np.float64))
    y, c = forward(p, x, mode)
    t = np.asarray(t, dtype=np.float64)
    m = np.asarray(m, dtype=np.float64)
    B = y.shape[0]
    e = y[:, :3] - t[:, :3]
    loss = np.sum(REG_W * m[:, :3] * e * e) / B
    z = y[:, 3:]
    wb = 0.1 / (2.0 * B) * m[:, 3:]
    loss += np.sum(wb * (softplus(z) - t[:, 3:] * z))
    dy = np.zeros_like(y)
    dy[:, :3] = 2.0 * REG_W * m[:, :3] * e / B
    dy[:, 3:] = wb * (sigmoid(z) - t[:, 3:])
    g = np.zeros(NPARAMS)
    g[W4:B4] = (dy.T @ r(c["a3"])).ravel()
    g[B4:] = dy.sum(0)
    dz3 = (dy @ r(q["W4"])) * dsilu(c["z3"])
    g[W3:B3] = (r(dz3).T @ r(c["a2"])).ravel()
    g[B3:W4] = dz3.sum(0)
    dz2 = (r(dz3) @ r(q["W3"])) * dsilu(c["z2"])
    g[W2:B2] = (r(dz2).T @ r(c["a1"])).ravel()
    g[B2:W3] = dz2.sum(0)
    dz1 = (r(dz2) @ r(q["W2"])) * dsilu(c["z1"])
    g[W1:B1] = (dz1.T @ c["x"]).ravel()
    g[B1:W2] = dz1.sum(0)
    return loss, g


def init_params(seed: int) -> np.ndarray:
    """A He-style random parameter vector (for tests; the library has its own initialiser)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS)
    p[W1:B1] = rng.uniform(-1, 1, H * NIN) * np.sqrt(3.0 / NIN) * 1.7
    p[W2:B2] = rng.uniform(-1, 1, H * H) * np.sqrt(3.0 / H) * 1.7
    p[W3:B3] = rng.uniform(-1, 1, H * H) * np.sqrt(3.0 / H) * 1.7
    p[W4:B4] = rng.uniform(-1, 1, NOUT * H) * np.sqrt(3.0 / H) * 0.17
    for a, b in ((B1, W2), (B2, W3), (B3, W4), (B4, NPARAMS)):
        p[a:b] = rng.uniform(-0.2, 0.2, b - a)
    return p
