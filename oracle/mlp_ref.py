"""CPU checker (test infrastructure only) for the learned guess's network kernels
(include/tonto_mlp.h): the same three-layer SiLU network, MSE loss and Adam step in
numpy float64.  Only tests import this; the product path runs the HIP kernels.

Pinned by tests/test_guess_cpu.py: the analytic gradient against central finite
differences, the Adam step against its closed form on a one-parameter problem.
"""
import numpy as np


def split(params, n_in, nh, n_out):
    p = np.asarray(params, np.float64)
    o = 0
    out = []
    for a, b in ((n_in, nh), (nh, nh), (nh, n_out)):
        W = p[o:o + a * b].reshape(a, b)
        o += a * b
        bb = p[o:o + b]
        o += b
        out += [W, bb]
    assert o == p.size
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, n_in, nh, n_out, X):
    W1, b1, W2, b2, W3, b3 = split(params, n_in, nh, n_out)
    X = np.asarray(X, np.float64)
    Z1 = X @ W1 + b1
    A1 = silu(Z1)
    Z2 = A1 @ W2 + b2
    A2 = silu(Z2)
    Y = A2 @ W3 + b3
    return Y, (X, Z1, A1, Z2, A2)


def backward(params, n_in, nh, n_out, cache, dY):
    W1, b1, W2, b2, W3, b3 = split(params, n_in, nh, n_out)
    X, Z1, A1, Z2, A2 = cache
    dY = np.asarray(dY, np.float64)
    gW3 = A2.T @ dY
    gb3 = dY.sum(0)
    dZ2 = (dY @ W3.T) * silu_grad(Z2)
    gW2 = A1.T @ dZ2
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2.T) * silu_grad(Z1)
    gW1 = X.T @ dZ1
    gb1 = dZ1.sum(0)
    return np.concatenate([g.ravel() for g in (gW1, gb1, gW2, gb2, gW3, gb3)])


def mse(Y, T, scale):
    d = np.asarray(Y, np.float64) - np.asarray(T, np.float64)
    return scale * float(np.sum(d * d)), 2.0 * scale * d


def adam(p, g, m, v, lr, b1, b2, eps, step):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
