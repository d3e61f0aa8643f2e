"""CPU checks of the split-BF16 product emulation (tests/split_emul.py), the
checker of the SKKR_MIXED_BF16 GEMM kernel."""
import numpy as np
import pytest

import split_emul as se


def _rand(rng, m, n, rowscale=True):
    a = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    if rowscale:  # rows spanning 12 orders of magnitude
        a *= np.logspace(-6, 6, m)[:, None]
    return a


@pytest.mark.parametrize("S", [2, 3, 4, 5])
@pytest.mark.parametrize("shape", [(64, 64, 64), (36, 36, 36), (128, 128, 128), (256, 128, 128), (20, 50, 70)])
def test_error_within_bound(S, shape):
    rng = np.random.default_rng(hash((S,) + shape) % 2**32)
    m, k, n = shape
    A = _rand(rng, m, k)
    B = _rand(rng, k, n).T.copy().T  # column scales via the transpose below
    B = B * np.logspace(3, -3, n)[None, :]
    C = se.split_product(A, B, S)
    ref = A @ B
    bnd = se.error_bound(A, B, S) + 1e-14 * (np.abs(A) @ np.abs(B))
    assert (np.abs(C.real - ref.real) <= bnd).all()
    assert (np.abs(C.imag - ref.imag) <= bnd).all()


def test_error_falls_with_slices():
    rng = np.random.default_rng(5)
    A, B = _rand(rng, 128, 128), _rand(rng, 128, 128, rowscale=False)
    ref = A @ B
    # normwise relative error against the row/column scale
    errs = []
    for S in (2, 3, 4, 5):
        C = se.split_product(A, B, S)
        errs.append(np.max(np.abs(C - ref) / (np.abs(A).max(1)[:, None] * np.abs(B).max(0)[None, :])))
    assert errs[0] > errs[1] > errs[2] > errs[3]
    # each extra slice buys about 9 bits
    for a, b in zip(errs, errs[1:]):
        assert 2 ** 6 < a / b < 2 ** 12
    assert errs[3] < 2.0 ** -38


def test_group_sums_stay_exact_in_fp32():
    """For S <= 4 every group sum is an integer below 2^24, even for the
    worst case (all slices at their maximum) -- the FP32 MFMA accumulation is
    then exact."""
    rng = np.random.default_rng(1)
    worst = np.full((64, 64), 0.999 + 0.999j)
    for S in (2, 3, 4):
        for A, B in ((worst, worst), (_rand(rng, 64, 64), _rand(rng, 64, 64))):
            _, gmax = se.split_product(A, B, S, return_groups=True)
            assert gmax <= 2.0 ** 24


def test_small_integers_exact():
    """Entries that the first slice holds exactly give the exact product."""
    rng = np.random.default_rng(2)
    A = rng.integers(-3, 4, (48, 64)) + 1j * rng.integers(-3, 4, (48, 64))
    B = rng.integers(-3, 4, (64, 40)) + 1j * rng.integers(-3, 4, (64, 40))
    for S in (2, 5):
        assert np.array_equal(se.split_product(A.astype(complex), B.astype(complex), S), A @ B)


def test_slices_recombine():
    rng = np.random.default_rng(3)
    y = rng.uniform(-1, 1, 1000)
    for S in (2, 4, 5):
        s = se.slices(y, S)
        assert np.abs(s).max() <= 256
        w = 2.0 ** (-8 - 9 * np.arange(S))
        assert np.abs((s * w[:, None]).sum(0) - y).max() <= 2.0 ** (-9 * S)


def test_zero_and_scale_invariance():
    rng = np.random.default_rng(4)
    A, B = _rand(rng, 64, 64), _rand(rng, 64, 64)
    A[3] = 0
    C = se.split_product(A, B, 4)
    assert (C[3] == 0).all()
    # power-of-two scaling of a row scales the result exactly
    A2 = A.copy()
    A2[5] *= 2.0 ** 40
    C2 = se.split_product(A2, B, 4)
    assert np.array_equal(C2[5], C[5] * 2.0 ** 40)
