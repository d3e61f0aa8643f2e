"""CPU checks of the INT8 digit product emulation (tests/i8_emul.py), the
checker of the SKKR_MIXED_INT8 GEMM kernel."""
import numpy as np
import pytest

import i8_emul as ie


def _rand(rng, m, n, rowscale=True):
    a = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    if rowscale:  # rows spanning 12 orders of magnitude
        a *= np.logspace(-6, 6, m)[:, None]
    return a


@pytest.mark.parametrize("S", [2, 3, 4, 5, 6])
def test_digits_represent_the_integer(S):
    """Balanced base-256 digits: sum d_t 256^t = V exactly, every digit a
    signed byte, and |V| <= 2^(8S-2) for values below 2^e."""
    rng = np.random.default_rng(S)
    x = rng.uniform(-1, 1, 5000) * 0.999999
    x[:4] = [0.0, -0.0, 0.99999999999, -0.99999999999]
    e = np.zeros_like(x, dtype=np.int64)
    V = ie.to_int(x, e, S)
    assert np.abs(V).max() <= 2 ** (8 * S - 2)
    d = ie.digits(V, S)
    assert d.min() >= -128 and d.max() <= 127
    back = sum(d[t].astype(object) * 256 ** t for t in range(S))
    assert all(int(b) == int(v) for b, v in zip(back, V))


@pytest.mark.parametrize("S", [2, 4, 6])
def test_magic_number_cut_equals_rint_digits(S):
    """The kernel cuts digits from the bits of x 2^(8S-2-e) + 1.5 2^52 + c;
    they equal the digits of rint(x 2^(8S-2-e)), ties to even included."""
    rng = np.random.default_rng(10 + S)
    x = rng.standard_normal(20000) * np.exp(rng.uniform(-30, 30, 20000))
    mx = np.abs(x)
    _, e = np.frexp(mx)
    e = e.astype(np.int64)
    # exact ties: x 2^(8S-2-e) = k + 1/2
    t = (np.arange(200) + 0.5) * 2.0 ** (-(8 * S - 2))
    x = np.concatenate([x, t, -t])
    e = np.concatenate([e, np.zeros(400, np.int64)])
    np.testing.assert_array_equal(ie.magic_digits(x, e, S), ie.digits(ie.to_int(x, e, S), S))


@pytest.mark.parametrize("S", [2, 3, 4, 5, 6])
@pytest.mark.parametrize("shape", [(64, 64, 64), (36, 36, 36), (128, 128, 128), (256, 128, 128), (20, 50, 70)])
def test_error_within_bound(S, shape):
    rng = np.random.default_rng(hash((S,) + shape) % 2**32)
    m, k, n = shape
    A = _rand(rng, m, k)
    B = _rand(rng, k, n) * np.logspace(3, -3, n)[None, :]
    C = ie.i8_product(A, B, S)
    ref = A @ B
    bnd = ie.error_bound(A, B, S) + 1e-14 * (np.abs(A) @ np.abs(B))
    assert (np.abs(C.real - ref.real) <= bnd).all()
    assert (np.abs(C.imag - ref.imag) <= bnd).all()


def test_error_falls_with_digits():
    rng = np.random.default_rng(5)
    A, B = _rand(rng, 128, 128), _rand(rng, 128, 128, rowscale=False)
    ref = A @ B
    errs = []
    for S in (2, 3, 4, 5, 6):
        C = ie.i8_product(A, B, S)
        errs.append(np.max(np.abs(C - ref) / (np.abs(A).max(1)[:, None] * np.abs(B).max(0)[None, :])))
    # each extra digit buys about 8 bits; six digits: about K 2^-46 (K = 128)
    for a, b in zip(errs, errs[1:]):
        assert 2 ** 5 < a / b < 2 ** 11
    assert errs[-1] < 2.0 ** -38


def test_group_pairs_stay_in_int32():
    """P_h = 256 G_2h + G_2h+1 stays below 2^31 even when every digit is at
    its extreme (values just below 2^e, and -2^e-ish), for every S."""
    rng = np.random.default_rng(1)
    worst = np.full((64, 64), 0.99999999 + 0.99999999j)
    alt = np.where(rng.random((64, 64)) < 0.5, -1, 1) * (0.99999999 - 0.99999999j)
    for S in (2, 3, 4, 5, 6):
        for A, B in ((worst, worst), (worst, alt), (_rand(rng, 64, 64), _rand(rng, 64, 64))):
            _, pmax = ie.i8_product(A, B, S, return_groups=True)
            assert pmax < 2.0 ** 31


def test_small_integers_exact():
    """Entries held exactly by the top digits give the exact product."""
    rng = np.random.default_rng(3)
    A = rng.integers(-50, 50, (40, 70)) + 1j * rng.integers(-50, 50, (40, 70))
    B = rng.integers(-50, 50, (70, 30)) + 1j * rng.integers(-50, 50, (70, 30))
    for S in (2, 4, 6):
        np.testing.assert_array_equal(ie.i8_product(A, B, S), A @ B)


def test_zero_lines():
    """All-zero rows and columns of a chunk give exact zeros; a matrix scaled
    by a power of two gives the product scaled by it (the exponents shift)."""
    rng = np.random.default_rng(4)
    A, B = _rand(rng, 48, 40), _rand(rng, 40, 24)
    A[5] = 0
    B[:, 7] = 0
    C = ie.i8_product(A, B, 6)
    assert (C[5] == 0).all() and (C[:, 7] == 0).all()
    np.testing.assert_array_equal(ie.i8_product(A * 2.0 ** 37, B * 2.0 ** -11, 6), C * 2.0 ** 26)
