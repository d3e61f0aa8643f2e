"""The error-prediction MLP's HIP kernels (libskkr_mlp.so) against the numpy
FP64 reference: forward, loss, gradient (the reference's gradient is pinned
by finite differences in test_mlp_oracle.py) and GPU training."""
import numpy as np
import pytest

from oracle import mlp_ref

pytestmark = pytest.mark.gpu


def _data(n=3000, nin=7, seed=5):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, nin)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] ** 2 - 0.3 * x[:, 2] * x[:, 3]).astype(np.float32)
    return x, t


def test_forward_loss_and_gradient_match_reference():
    from skkr_budapest_amd.mlp import Mlp
    x, t = _data(n=1100)            # three gradient chunks, the last ragged
    m = Mlp(x.shape[1], seed=11)
    p = m.params().astype(np.float64)
    print("mlp: n=1100 nin=7 forward/loss/grad", flush=True)
    y = m.predict(x)
    yr, _, _ = mlp_ref.forward(p, x.astype(np.float64), x.shape[1])
    assert np.max(np.abs(y - yr)) < 1e-5 * max(1.0, np.max(np.abs(yr)))
    loss, g = m.loss_grad(x, t)
    lr_ = mlp_ref.loss(p, x.astype(np.float64), t.astype(np.float64), x.shape[1])
    gr = mlp_ref.grad(p, x.astype(np.float64), t.astype(np.float64), x.shape[1])
    assert abs(loss - lr_) < 1e-5 * lr_
    # FP32 sums over 1100 samples: 1e-4 of the gradient norm
    assert np.linalg.norm(g - gr) < 1e-4 * np.linalg.norm(gr)
    loss2, g2 = m.loss_grad(x, t)
    assert loss2 == loss and np.array_equal(g, g2)      # fixed summation order


def test_set_params_round_trip_and_single_sample():
    from skkr_budapest_amd.mlp import Mlp
    x, t = _data(n=1)
    m = Mlp(x.shape[1], seed=2)
    p = np.random.default_rng(1).normal(scale=0.1, size=m.np).astype(np.float32)
    m.set_params(p)
    assert np.array_equal(m.params(), p)
    loss, g = m.loss_grad(x, t)
    gr = mlp_ref.grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), x.shape[1])
    assert np.linalg.norm(g - gr) < 1e-4 * np.linalg.norm(gr)


def test_training_on_gpu_reduces_the_loss():
    from skkr_budapest_amd.mlp import Mlp
    x, t = _data()
    m = Mlp(x.shape[1], seed=3)
    print("mlp: training 600 Adam steps on 3000 samples", flush=True)
    hist = m.fit(x, t, steps=600, lr=3e-3)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.05 * hist[0]
    # the trained network generalises to fresh samples of the same law
    xv, tv = _data(n=2000, seed=9)
    assert np.mean((m.predict(xv) - tv) ** 2) < 0.1 * np.var(tv)
