"""CPU reference of the surrogate MLP (6 -> 64 -> 64 -> 1, GELU).

TEST INFRASTRUCTURE ONLY: the tests check the HIP training and inference kernels
(sparkx_amd/csrc/hb_mlp.hip) against it.  No reference counterpart: the reference has no
learned component; the model is the surrogate network those kernels train.

Arithmetic of the GPU model, restated here with float64 accumulation:
    xn = (x - mean) / std                       (f32)
    a0 = bf16(xn)
    z1 = a0 @ bf16(W1) + b1;   h1 = gelu(z1);   a1 = bf16(h1)
    z2 = a1 @ bf16(W2) + b2;   h2 = gelu(z2);   a2 = bf16(h2)
    y  = a2 @ bf16(W3) + b3
    loss = mean((y - t)^2)
Backward (the GPU rounds the matmul operands dZ to bf16 the same way):
    dy = 2 (y - t) / B
    gW3 = a2^T dy,  gb3 = sum dy,   dh2 = dy bf16(W3)^T
    dz2 = dh2 * gelu'(z2),  d2 = bf16(dz2),  gW2 = a1^T d2,  gb2 = sum dz2
    dh1 = d2 bf16(W2)^T,    dz1 = dh1 * gelu'(z1), d1 = bf16(dz1),
    gW1 = a0^T d1,  gb1 = sum dz1
gelu(x) = x Phi(x), gelu'(x) = Phi(x) + x phi(x) (exact, erf-based).
"""
from __future__ import annotations

import math

import numpy as np

D_IN, D_H = 6, 64
# flat parameter layout (f32 master copy on the GPU)
W1, B1, W2, B2, W3, B3 = (0, 384, 448, 4544, 4608, 4672)
N_PARAMS = 4673


def split(params):
    p = np.asarray(params, dtype=np.float64)
    return (p[W1:B1].reshape(D_IN, D_H), p[B1:W2], p[W2:B2].reshape(D_H, D_H), p[B2:W3], p[W3:B3], p[B3])


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), returned as float64."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


_erf = np.vectorize(math.erf)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    cdf = 0.5 * (1.0 + _erf(x / math.sqrt(2.0)))
    pdf = np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)
    return cdf + x * pdf


def normalize(x, mean, std):
    x = np.asarray(x, np.float32)
    return ((x - np.asarray(mean, np.float32)) / np.asarray(std, np.float32)).astype(np.float32)


def forward(params, x, mean, std, emulate_bf16=True):
    w1, b1, w2, b2, w3, b3 = split(params)
    r = bf16 if emulate_bf16 else (lambda v: np.asarray(v, np.float64))
    a0 = r(normalize(x, mean, std))
    z1 = a0 @ r(w1) + b1
    a1 = r(gelu(z1))
    z2 = a1 @ r(w2) + b2
    a2 = r(gelu(z2))
    y = a2 @ r(w3) + b3
    return y, (a0, z1, a1, z2, a2)


def loss_and_grad(params, x, t, mean, std, emulate_bf16=True):
    w1, b1, w2, b2, w3, b3 = split(params)
    r = bf16 if emulate_bf16 else (lambda v: np.asarray(v, np.float64))
    y, (a0, z1, a1, z2, a2) = forward(params, x, mean, std, emulate_bf16)
    t = np.asarray(t, np.float64)
    B = y.shape[0]
    loss = float(np.mean((y - t) ** 2))
    dy = 2.0 * (y - t) / B
    g = np.zeros(N_PARAMS)
    g[W3:B3] = a2.T @ dy
    g[B3] = dy.sum()
    dh2 = np.outer(dy, r(w3))
    dz2 = dh2 * gelu_grad(z2)
    d2 = r(dz2)
    g[W2:B2] = (a1.T @ d2).ravel()
    g[B2:W3] = dz2.sum(axis=0)
    dh1 = d2 @ r(w2).T
    dz1 = dh1 * gelu_grad(z1)
    d1 = r(dz1)
    g[W1:B1] = (a0.T @ d1).ravel()
    g[B1:W2] = dz1.sum(axis=0)
    return loss, g


def finite_difference(params, x, t, mean, std, idx, eps=1e-6):
    """Central differences of the float64 loss (no bf16 rounding) at parameters idx."""
    out = []
    for i in idx:
        p = np.array(params, dtype=np.float64)
        p[i] += eps
        lp = float(np.mean((forward(p, x, mean, std, False)[0] - t) ** 2))
        p[i] -= 2 * eps
        lm = float(np.mean((forward(p, x, mean, std, False)[0] - t) ** 2))
        out.append((lp - lm) / (2 * eps))
    return np.array(out)


def adam(params, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) in float64; returns new params, m, v."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
