"""Numpy FP64 reference of the error-prediction MLP (TEST INFRASTRUCTURE):
the same architecture and parameter layout as skkr_budapest_amd/csrc/mlp.hip
(W1[nh][nin], b1, W2[nh][nh], b2, w3[nh], b3), its mean-squared-error loss and
the analytic gradient, pinned in the tests by central finite differences."""
import numpy as np

NH = 64


def unpack(p, nin, nh=NH):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + nh * nin].reshape(nh, nin); o += nh * nin
    b1 = p[o:o + nh]; o += nh
    W2 = p[o:o + nh * nh].reshape(nh, nh); o += nh * nh
    b2 = p[o:o + nh]; o += nh
    w3 = p[o:o + nh]; o += nh
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, x, nin):
    W1, b1, W2, b2, w3, b3 = unpack(p, nin)
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ w3 + b3, h1, h2


def loss(p, x, t, nin):
    y, _, _ = forward(p, x, nin)
    return float(np.mean((y - t) ** 2))


def grad(p, x, t, nin):
    W1, b1, W2, b2, w3, b3 = unpack(p, nin)
    y, h1, h2 = forward(p, x, nin)
    n = len(x)
    d3 = 2.0 * (y - t) / n
    d2 = (d3[:, None] * w3[None, :]) * (1 - h2 ** 2)
    d1 = (d2 @ W2) * (1 - h1 ** 2)
    return np.concatenate([(d1.T @ x).ravel(), d1.sum(0), (d2.T @ h1).ravel(), d2.sum(0),
                           d3 @ h2, [d3.sum()]])
