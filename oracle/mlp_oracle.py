"""CPU oracle of the tuner's cost-model perceptron (test infrastructure only:
tests use it to check the HIP forward, backward and Adam kernels; the runtime
never calls it).

The reference has no learned cost model (its tuner measures block shapes
online, rts/dvmh/loop.cpp:1487-1909), so there is no reference code to
follow: this restates the mathematics of the network the runtime implements,
in float64.

Network: x (F) -> ReLU(W1 x + b1) (H) -> ReLU(W2 h1 + b2) (H) -> W3 . h2 + b3.
Parameter vector: W1[H][F] | b1[H] | W2[H][H] | b2[H] | W3[H] | b3.
Loss: mean over samples of (out - y)^2.  ReLU'(0) is taken as 0.

Pinned by: central finite differences of the loss (gradient), a network whose
ReLUs are all active reducing to an affine map with a closed-form gradient,
and Adam's first step, which moves every parameter by lr * g / (|g| + eps')
(Kingma & Ba 2015, Algorithm 1).
"""
import numpy as np

F = 24
H = 64
NPARAM = H * F + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3, [b3]])


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64).reshape(-1, F)
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    return h2 @ W3 + b3


def loss(p, X, y):
    e = forward(p, X) - np.asarray(y, dtype=np.float64)
    return float(np.mean(e * e))


def grad(p, X, y):
    """Gradient of loss() with respect to the parameter vector, and the loss."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64).reshape(-1, F)
    y = np.asarray(y, dtype=np.float64)
    n = X.shape[0]
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    out = h2 @ W3 + b3
    e = out - y
    dout = 2.0 * e / n
    gb3 = dout.sum()
    gW3 = h2.T @ dout
    d2 = np.outer(dout, W3) * (a2 > 0)
    gb2 = d2.sum(axis=0)
    gW2 = d2.T @ h1
    d1 = (d2 @ W2) * (a1 > 0)
    gb1 = d1.sum(axis=0)
    gW1 = d1.T @ X
    return pack(gW1, gb1, gW2, gb2, gW3, gb3), float(np.mean(e * e))


def adam_step(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (step counts from 1).  Returns new (p, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(rng):
    """Random parameters of the scale the runtime's initialisation uses."""
    W1 = rng.uniform(-1, 1, (H, F)) * np.sqrt(6.0 / F)
    W2 = rng.uniform(-1, 1, (H, H)) * np.sqrt(6.0 / H)
    W3 = rng.uniform(-1, 1, H) * np.sqrt(3.0 / H)
    return pack(W1, rng.normal(0, 0.1, H), W2, rng.normal(0, 0.1, H), W3, 0.05)
