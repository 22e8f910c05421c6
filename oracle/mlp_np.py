"""FP64 numpy restatement of the near-pair surrogate network.

TEST INFRASTRUCTURE ONLY: the checker that the HIP network kernels
(efie_mlp_forward, efie_mlp_loss_grad) are compared against; the product
never imports it.  The network has no counterpart in the reference.

Network: 13 -> 64 -> 64 -> 64 -> 16, exact GELU (x Phi(x)) after each
hidden layer, loss sum((y - t)^2) / (16 N).  Parameters are one flat array
W0[64][13] b0[64] W1[64][64] b1[64] W2[64][64] b2[64] W3[16][64] b3[16]
(row-major W[out][in]).

Pinned by tests/test_oracle_mlp.py: the gradient against central finite
differences of the loss, GELU' against finite differences of GELU, and the
forward pass of a one-hidden-unit special case against a hand computation.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

SHAPES = [(64, 13), (64, 64), (64, 64), (16, 64)]
N_PARAMS = sum(o * i + o for o, i in SHAPES)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    out, off = [], 0
    for o, i in SHAPES:
        W = p[off:off + o * i].reshape(o, i)
        off += o * i
        b = p[off:off + o]
        off += o
        out.append((W, b))
    return out


def forward(params, X):
    """Network output Y [N][16] for inputs X [N][13]."""
    h = np.asarray(X, dtype=np.float64)
    layers = unpack(params)
    for li, (W, b) in enumerate(layers):
        z = h @ W.T + b
        h = gelu(z) if li < len(layers) - 1 else z
    return h


def loss_grad(params, X, T):
    """(loss, flat gradient) of sum((y - t)^2) / (16 N)."""
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    N = X.shape[0]
    layers = unpack(params)
    hs, zs = [X], []
    h = X
    for li, (W, b) in enumerate(layers):
        z = h @ W.T + b
        zs.append(z)
        h = gelu(z) if li < len(layers) - 1 else z
        hs.append(h)
    y = hs[-1]
    e = y - T
    loss = float(np.sum(e * e) / (16.0 * N))
    d = 2.0 * e / (16.0 * N)
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        grads[li] = (d.T @ hs[li], d.sum(0))
        if li > 0:
            d = (d @ W) * gelu_d(zs[li - 1])
    flat = np.concatenate([np.concatenate([g.ravel(), gb]) for g, gb in grads])
    return loss, flat


def glorot(rng):
    """Random parameters: Glorot-uniform weights, small random biases."""
    parts = []
    for o, i in SHAPES:
        a = np.sqrt(6.0 / (o + i))
        parts.append(rng.uniform(-a, a, o * i))
        parts.append(rng.uniform(-0.1, 0.1, o))
    return np.concatenate(parts)
