"""Oracle for the learned-radius network -- TEST INFRASTRUCTURE ONLY.

The reference has no learned component; this is the mathematics of the
device network (F -> 64 -> 64 -> 1, ReLU, mean squared error, Adam with
beta1 = 0.9, beta2 = 0.999, eps = 1e-8) in float64 numpy, parameters laid
out as the library's: W1[64][F], b1[64], W2[64][64], b2[64], W3[64], b3.

Pinned by tests/test_mlp_oracle.py against central finite differences and
against torch autograd + torch.optim.Adam in float64 (independent
implementations), so a dropped term or a transposed weight fails there.
"""
from __future__ import annotations

import numpy as np

H = 64


def param_count(F: int) -> int:
    return H * F + H + H * H + H + H + 1


def unpack(p, F):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(p, x):
    F = x.shape[1]
    W1, b1, W2, b2, W3, b3 = unpack(p, F)
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    return h2 @ W3 + b3, (z1, h1, z2, h2)


def loss_and_grad(p, x, t):
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    B, F = x.shape
    W1, b1, W2, b2, W3, b3 = unpack(p, F)
    y, (z1, h1, z2, h2) = forward(p, x)
    loss = np.mean((y - t) ** 2)
    d3 = 2.0 * (y - t) / B
    d2 = np.outer(d3, W3) * (z2 > 0)
    d1 = (d2 @ W2) * (z1 > 0)
    g = np.concatenate([(d1.T @ x).ravel(), d1.sum(0), (d2.T @ h1).ravel(), d2.sum(0), h2.T @ d3, [d3.sum()]])
    return loss, g, y


def adam_first_step(p, g, lr):
    """One Adam step from zero moments (step 1)."""
    b1, b2, eps = 0.9, 0.999, 1e-8
    m = (1 - b1) * g
    v = (1 - b2) * g * g
    mh = m / (1 - b1)
    vh = v / (1 - b2)
    return np.asarray(p, np.float64) - lr * mh / (np.sqrt(vh) + eps)
