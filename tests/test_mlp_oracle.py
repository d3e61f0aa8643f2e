"""The numpy MLP reference (oracle/mlp_ref.py) pinned by central finite
differences of its loss, and the parameter layout of include/skkr_mlp.h."""
import numpy as np

from oracle import mlp_ref


def _setup(nin=7, n=40, seed=3):
    rng = np.random.default_rng(seed)
    npar = mlp_ref.NH * nin + mlp_ref.NH * mlp_ref.NH + 3 * mlp_ref.NH + 1
    p = rng.normal(scale=0.2, size=npar)
    x = rng.normal(size=(n, nin))
    t = rng.normal(size=n)
    return p, x, t, nin


def test_parameter_count_matches_header_layout():
    p, x, t, nin = _setup()
    W1, b1, W2, b2, w3, b3 = mlp_ref.unpack(p, nin)
    assert W1.shape == (64, nin) and W2.shape == (64, 64) and w3.shape == (64,)
    assert len(p) == 64 * nin + 64 + 64 * 64 + 64 + 64 + 1


def test_gradient_matches_central_differences():
    p, x, t, nin = _setup()
    g = mlp_ref.grad(p, x, t, nin)
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.choice(len(p), 60, replace=False), [len(p) - 1]])
    h = 1e-6
    for q in idx:
        e = np.zeros_like(p)
        e[q] = h
        fd = (mlp_ref.loss(p + e, x, t, nin) - mlp_ref.loss(p - e, x, t, nin)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-7 * max(1.0, abs(fd)), (q, fd, g[q])


def test_gradient_of_a_planted_output_bias_error():
    """With all weights zero the network outputs b3, so the loss is
    mean((b3 - t)^2) and only d/db3 = 2 (b3 - mean t) is nonzero."""
    p, x, t, nin = _setup()
    p[:] = 0.0
    p[-1] = 0.7
    g = mlp_ref.grad(p, x, t, nin)
    assert np.allclose(g[:-1], 0.0)
    assert np.isclose(g[-1], 2 * (0.7 - t.mean()))
    assert np.isclose(mlp_ref.loss(p, x, t, nin), np.mean((0.7 - t) ** 2))
