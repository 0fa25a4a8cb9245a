"""CPU oracle of the learned-jump MLP (FP64 numpy).  TEST INFRASTRUCTURE ONLY.

The MLP has no counterpart in the reference.  This restates the network the
HIP kernels implement (egg-sync_amd/csrc/egg_mlp.hip): in -> h -> h -> out,
ReLU, mean-over-batch squared error, flat parameter layout
W1[h][in], b1[h], W2[h][h], b2[h], W3[out][h], b3[out].  It is pinned by a
central finite-difference check (tests/test_oracle.py), and in turn checks the
GPU forward pass and gradients (tests/test_gpu_learned.py).
"""
from __future__ import annotations

import numpy as np


def sizes(n_in: int, h: int, n_out: int):
    return [("W1", (h, n_in)), ("b1", (h,)), ("W2", (h, h)), ("b2", (h,)), ("W3", (n_out, h)),
            ("b3", (n_out,))]


def unpack(P, n_in, h, n_out):
    P = np.asarray(P, np.float64)
    out, off = {}, 0
    for name, shape in sizes(n_in, h, n_out):
        k = int(np.prod(shape))
        out[name] = P[off:off + k].reshape(shape)
        off += k
    assert off == P.size
    return out


def forward(P, X, n_in, h, n_out):
    p = unpack(P, n_in, h, n_out)
    X = np.asarray(X, np.float64)
    z1 = X @ p["W1"].T + p["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ p["W2"].T + p["b2"]
    h2 = np.maximum(z2, 0.0)
    return h2 @ p["W3"].T + p["b3"], (X, z1, h1, z2, h2)


def loss_and_grad(P, X, T, n_in, h, n_out):
    """Loss mean_b sum_o (O - T)^2 and its gradient (flat, parameter layout)."""
    p = unpack(P, n_in, h, n_out)
    O, (X, z1, h1, z2, h2) = forward(P, X, n_in, h, n_out)
    B = X.shape[0]
    R = O - np.asarray(T, np.float64)
    loss = float(np.sum(R * R) / B)
    dO = 2.0 * R / B
    gW3 = dO.T @ h2
    gb3 = dO.sum(0)
    dh2 = (dO @ p["W3"]) * (z2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ p["W2"]) * (z1 > 0)
    gW1 = dh1.T @ X
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def adam_step(P, G, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * G
    v = b2 * v + (1 - b2) * G * G
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return P - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
