"""Inversions whose 16 x 16 diagonal blocks cannot be pivoted within
themselves: the blocked kernel flags them (zero pivot or multiplier growth)
and they are inverted again with the row-wise kernel, which pivots over all
rows as the reference's gjinv does (skkr_kpar_v6/main/gjinv.f:13-22).

Tolerance: relative Frobenius error <= 1e-12 against numpy's LU inverse for
the well-conditioned test matrices (condition numbers ~10)."""
import dataclasses

import numpy as np
import pytest

import oracle
from skkr_budapest_amd import systems as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from skkr_budapest_amd import gpu as G
    return G


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _well_conditioned(rng, n):
    a = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(2 * n)
    return a + 2.0 * np.eye(n)


@pytest.mark.parametrize("n", [32, 64, 128, 200])
def test_zero_leading_block(gpu, n):
    """Leading 16 x 16 block zero (and, for n = 32, the exact block
    permutation [[0, I], [I, 0]]): no pivot inside the block."""
    rng = np.random.default_rng(n)
    mats = []
    for j in range(6):
        a = _well_conditioned(rng, n)
        a[:16, :16] = 0.0
        mats.append(a)
    if n == 32:
        p = np.zeros((32, 32), complex)
        p[:16, 16:] = np.eye(16)
        p[16:, :16] = np.eye(16)
        mats.append(p)
    a = np.array(mats)
    with gpu.Context(lk=n, sites_per_pl=1, n_pl=1) as ctx:
        inv, st = ctx.gjinv(a)
    assert (st == 0).all()
    for j in range(len(a)):
        assert rel(inv[j], np.linalg.inv(a[j])) < 1e-12


@pytest.mark.parametrize("n", [64, 128])
def test_nearly_singular_leading_block(gpu, n):
    """Leading block 1e-9 times a random matrix inside a well-conditioned
    matrix: restricted pivoting would divide by ~1e-9 (multipliers ~1e9); the
    flagged matrices come back as accurate as with full pivoting, the others
    are untouched."""
    rng = np.random.default_rng(7 + n)
    a = np.array([_well_conditioned(rng, n) for _ in range(5)])
    a[1, :16, :16] *= 1e-9
    a[3, :16, :16] *= 1e-9
    with gpu.Context(lk=n, sites_per_pl=1, n_pl=1) as ctx:
        inv, st = ctx.gjinv(a)
        ctx.set_option("inv_kernel", 1)
        inv_rw, _ = ctx.gjinv(a)
    assert (st == 0).all()
    for j in range(5):
        assert rel(inv[j], np.linalg.inv(a[j])) < 1e-12
    # the flagged ones are exactly the row-wise kernel's result
    np.testing.assert_array_equal(inv[1], inv_rw[1])
    np.testing.assert_array_equal(inv[3], inv_rw[3])
    assert not np.array_equal(inv[0], inv_rw[0])


def test_singular_matrix_still_reported(gpu):
    rng = np.random.default_rng(3)
    a = _well_conditioned(rng, 64)
    a[:, 40] = 0.0  # a zero column: every pivot candidate of that column is exactly zero
    with gpu.Context(lk=64, sites_per_pl=1, n_pl=1) as ctx:
        _, st = ctx.gjinv(a[None])
    assert st[0] == 4  # SKKR_ERR_SINGULAR


def test_tau2d_with_unpivotable_diagonal_block(gpu):
    """tau2d (the sweeps) on blocks whose first PL matrix has a zero leading
    16 x 16 block: the sweeps run again with full pivoting and agree with the
    oracle's tau2d."""
    cfg = S.small_config(lk=16, s=2, P=3, r2max=2, energies=(0.3 + 0.1j,))
    sy = S.make_system(cfg)
    E = cfg.energies[0]
    n = cfg.n
    k, w = S.uniform_kmesh(2)
    tint, tl, tr = sy.tinv(E)
    rng = np.random.default_rng(11)
    nb = len(w)
    m00 = np.zeros((nb, n, n), complex)
    m01 = np.zeros((nb, n, n), complex)
    m10 = np.zeros((nb, n, n), complex)
    x = np.zeros((nb, n, n), complex)
    y = np.zeros((nb, n, n), complex)
    for j in range(nb):
        m00[j] = _well_conditioned(rng, n) * 3.0
        # A_11 = m00 + t^-1 - y: zero leading 16 x 16 block
        m00[j][:16, :16] = -tint[0]
        m01[j] = 0.1 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
        m10[j] = m01[j].conj().T
        x[j] = 0.05 * np.eye(n)
    with gpu.Context(lk=16, sites_per_pl=2, n_pl=3) as ctx:
        ctx.set_tinv(tint, tl, tr)
        tau, st = ctx.tau2d(m00, m01, m10, x, y)
    assert (st == 0).all()
    for j in range(nb):
        ref, rst = oracle.tau2d(m00[j], m01[j], m10[j], tint, 3, x[j], y[j])
        assert rst == 0
        assert rel(tau[j], ref) < 1e-12


def test_config1_literal_dense_tinv_law(gpu):
    """Config 1 with its t^-1 as the configuration text words it: dense
    random complex site blocks (every spin sub-block filled), diagonally
    dominant, no Hermitian structure -- 8 k-points against the oracle."""
    cfg = dataclasses.replace(S.CONFIGS[1], tmodel="dense", name="cfg1-dense")
    sy = S.make_system(cfg)
    E = cfg.energies[0]
    k, w = S.uniform_kmesh(cfg.mesh)
    idx = np.array([0, 517, 1300, 2047, 2048, 3001, 3500, 4095])
    ref = oracle.tau_bz(sy, E, k[idx], w[idx], nworkers=8)
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        r = ctx.tau_bz(E, k[idx], w[idx])
    np.testing.assert_array_equal(r["iters"], ref["iters"])
    assert rel(r["tau_bz"], ref["tau_bz"]) < 1e-12
