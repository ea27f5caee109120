"""CPU model of the learned-CDF network (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module.  The network has no counterpart in the reference (the reference
sorts with RadixSort<u64>, ops/sort.hpp:83-160); it is the 1 -> 32 -> 32 -> 1
ReLU MLP whose forward pass, mean-squared-error gradient and Adam step the GPU
library computes in FP32.  Here the same functions are written in float64
numpy with the parameter packing of the C ABI:

    W1[32] b1[32] W2[32][32] (row = output unit) b2[32] W3[32] b3

Pinning: tests/test_oracle.py checks ``grad`` against central finite
differences of ``loss`` and ``forward`` against a loop over units, so a
dropped term, a transposed W2 or a wrong ReLU mask fails there.
"""
from __future__ import annotations

import numpy as np

H = 32
N_PARAMS = 32 + 32 + 32 * 32 + 32 + 32 + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    assert p.size == N_PARAMS
    w1, b1 = p[0:32], p[32:64]
    w2 = p[64:64 + 1024].reshape(H, H)
    b2 = p[1088:1120]
    w3, b3 = p[1120:1152], p[1152]
    return w1, b1, w2, b2, w3, b3


def forward(p, x):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(np.outer(x, w1) + b1, 0.0)
    h2 = np.maximum(h1 @ w2.T + b2, 0.0)
    return h2 @ w3 + b3


def loss(p, x, y):
    e = forward(p, x) - np.asarray(y, np.float64)
    return float(np.mean(e * e))


def grad(p, x, y):
    """d loss / d params, packed like the parameters."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, np.float64)
    n = x.size
    a1 = np.outer(x, w1) + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ w2.T + b2
    h2 = np.maximum(a2, 0.0)
    yp = h2 @ w3 + b3
    dy = 2.0 * (yp - np.asarray(y, np.float64)) / n
    dh2 = np.where(a2 > 0, np.outer(dy, w3), 0.0)
    dh1 = np.where(a1 > 0, dh2 @ w2, 0.0)
    g = np.zeros(N_PARAMS)
    g[0:32] = dh1.T @ x
    g[32:64] = dh1.sum(0)
    g[64:1088] = (dh2.T @ h1).ravel()
    g[1088:1120] = dh2.sum(0)
    g[1120:1152] = h2.T @ dy
    g[1152] = dy.sum()
    return g


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (bias-corrected), returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
