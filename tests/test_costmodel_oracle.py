"""Pins of the cost-model oracle (oracle/mlp_oracle.py) on the CPU: the
gradient against central finite differences and against the closed form of
an all-active (affine) network, Adam's first step against its closed form."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'oracle'))
import mlp_oracle as M  # noqa: E402


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(11)
    p = M.init_params(rng)
    X = rng.normal(size=(37, M.F))
    y = rng.normal(size=37)
    g, l0 = M.grad(p, X, y)
    assert abs(l0 - M.loss(p, X, y)) < 1e-12
    idx = np.concatenate([rng.choice(M.NPARAM, 300, replace=False), [M.NPARAM - 1]])
    h = 1e-6
    for i in idx:
        e = np.zeros(M.NPARAM)
        e[i] = h
        fd = (M.loss(p + e, X, y) - M.loss(p - e, X, y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_gradient_of_affine_network_closed_form():
    """With positive weights, positive inputs and positive biases every ReLU is
    active, so out = W3 (W2 (W1 x + b1) + b2) + b3 is affine in x and the
    gradient of b3 is mean(2 e), of W3 is mean(2 e h2) with h2 = W2 h1 + b2."""
    rng = np.random.default_rng(3)
    W1 = rng.uniform(0.01, 0.1, (M.H, M.F))
    W2 = rng.uniform(0.01, 0.1, (M.H, M.H))
    W3 = rng.uniform(0.01, 0.1, M.H)
    b1 = rng.uniform(0.1, 0.2, M.H)
    b2 = rng.uniform(0.1, 0.2, M.H)
    p = M.pack(W1, b1, W2, b2, W3, 0.3)
    X = rng.uniform(0.0, 1.0, (9, M.F))
    y = rng.normal(size=9)
    h1 = X @ W1.T + b1
    h2 = h1 @ W2.T + b2
    out = h2 @ W3 + 0.3
    e = out - y
    g, _ = M.grad(p, X, y)
    _, _, _, _, gW3, gb3 = M.unpack(g)
    assert np.allclose(gb3, np.mean(2 * e), rtol=1e-12)
    assert np.allclose(gW3, (2 * e / len(y)) @ h2, rtol=1e-12)
    # d out / d b1 = W2^T W3 for every sample
    _, gb1, _, _, _, _ = M.unpack(g)
    assert np.allclose(gb1, np.sum(2 * e / len(y)) * (W2.T @ W3), rtol=1e-10)


def test_adam_first_step_closed_form():
    """At step 1 with zero moments, m_hat = g and v_hat = g^2, so the update is
    lr * g / (|g| + eps): lr * sign(g) up to eps."""
    rng = np.random.default_rng(5)
    p = rng.normal(size=M.NPARAM)
    g = rng.normal(size=M.NPARAM)
    z = np.zeros(M.NPARAM)
    p1, m1, v1 = M.adam_step(p, g, z, z, 1, 0.01)
    assert np.allclose(p - p1, 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)
    assert np.allclose(m1, 0.1 * g) and np.allclose(v1, 0.001 * g * g)
