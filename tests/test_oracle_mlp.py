"""Pins of the FP64 network oracle (oracle/mlp_np.py), the checker of the
HIP surrogate network kernels.

- GELU' against central differences of GELU;
- a network whose only nonzero path is one hidden unit per layer, against
  the same composition written out by hand;
- the full gradient against central finite differences of the loss
  (every parameter block, including biases and the 13-wide first layer).
"""
import numpy as np

import mlp_np as M


def test_gelu_derivative():
    z = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_d(z))) < 1e-9
    # GELU(0) = 0, GELU(x) -> x for large x, -> 0 for very negative x
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert abs(M.gelu(np.array([8.0]))[0] - 8.0) < 1e-12
    assert abs(M.gelu(np.array([-8.0]))[0]) < 1e-12


def test_single_path_by_hand():
    p = np.zeros(M.N_PARAMS)
    (W0, b0), (W1, b1), (W2, b2), (W3, b3) = M.unpack(p)
    # the views share memory with p
    W0[5, 2], b0[5] = 0.7, -0.1
    W1[9, 5], b1[9] = 1.3, 0.2
    W2[0, 9], b2[0] = -0.4, 0.05
    W3[3, 0], b3[3] = 2.0, 0.5
    x = np.zeros((1, 13))
    x[0, 2] = 0.9
    g = lambda v: 0.5 * v * (1 + __import__("math").erf(v / np.sqrt(2)))
    h0 = g(0.7 * 0.9 - 0.1)
    h1 = g(1.3 * h0 + 0.2)
    h2 = g(-0.4 * h1 + 0.05)
    y = M.forward(p, x)
    expect = np.zeros(16)
    expect[3] = 2.0 * h2 + 0.5
    # the other outputs see only the hidden units at GELU(0) = 0 and zero biases
    assert np.allclose(y[0], expect, atol=1e-15, rtol=1e-14)


def test_gradient_vs_finite_differences():
    rng = np.random.default_rng(3)
    p = M.glorot(rng)
    X = rng.normal(size=(7, 13))
    T = rng.normal(size=(7, 16))
    _, g = M.loss_grad(p, X, T)
    # sample indices from every block
    offs = np.cumsum([0] + [o * i + o for o, i in M.SHAPES])
    idx = []
    for k in range(4):
        idx += list(rng.integers(offs[k], offs[k + 1], 12))
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_grad(pp, X, T)[0] - M.loss_grad(pm, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])
    # a dropped bias term or a wrong sign in the output gradient fails this
    assert np.linalg.norm(g) > 0
