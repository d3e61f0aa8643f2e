"""tau2d of a layered stack (bulkgeo false) on the GPU, through the C-ABI
(skkr_tau2d_layered_batch), against the reference's own tau2d
(tests/golden/layered16.npz) and the oracle.  FP64 tolerance: relative
Frobenius error <= 1e-11."""
import os

import numpy as np
import pytest

import oracle
from test_layered_oracle import LG, layered_blocks

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TOL = 1e-11


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def gpu():
    from skkr_budapest_amd import gpu as G
    return G


def _inputs(sy, pls, k, E):
    ti, tl, tr = sy.tinv(E)
    L = [layered_blocks(sy, pls, kk) for kk in k]
    m00 = np.stack([b[1] for b in L]); m01 = np.stack([b[2] for b in L])
    m10 = np.stack([b[3] for b in L])
    Y = np.stack([oracle.dugo(*b[0], tl, "L")[0] for b in L])
    X = np.stack([oracle.dugo(*b[0], tr, "R")[0] for b in L])
    return (ti, tl, tr), m00, m01, m10, X, Y


def test_layered_tau2d_matches_reference_golden(gpu):
    g = np.load(os.path.join(GOLD, "layered16.npz"))
    sy, pls = LG.layered_system()
    E = complex(g["E"])
    k, w = g["k"], g["w"]
    (ti, tl, tr), m00, m01, m10, X, Y = _inputs(sy, pls, k, E)
    print("layered tau2d: lk 16, s 4, P 4, nk", len(w), flush=True)
    with gpu.Context.for_system(sy, batch=5) as ctx:
        ctx.set_tinv(ti, tl, tr)
        tau, st = ctx.tau2d_layered(m00, m01, m10, X, Y)
    assert (st == 0).all()
    for j, ik in enumerate(g["save_idx"]):
        assert rel(tau[ik], g["tau_save"][j]) < TOL
    assert rel(np.tensordot(w, tau, axes=(0, 0)), g["tau_bz"]) < TOL
    for ik in (0, 7, 15):
        ref, _ = oracle.tau2d_layered(m00[ik], m01[ik], m10[ik], ti, X[ik], Y[ik])
        assert rel(tau[ik], ref) < TOL


def test_layered_with_equal_planes_is_bulk_and_leaves_bulk_calls_alone(gpu):
    """Equal planes give the bulk tau2d bitwise (same kernels, same order); a
    bulk call after a layered one is unchanged (the layered blocks are only
    read inside the layered call)."""
    sy, pls = LG.layered_system()
    E = sy.cfg.energies[0]
    ti, tl, tr = sy.tinv(E)
    k = np.array([[0.3, -0.2], [1.7, 0.9], [-2.0, 0.4]])
    B = [oracle.pl_blocks(*LG.gk(sy, kk)) for kk in k]
    m00 = np.stack([b[0] for b in B]); m01 = np.stack([b[1] for b in B])
    m10 = np.stack([b[2] for b in B])
    X = np.stack([oracle.dugo(*b, tr, "R")[0] for b in B])
    Y = np.stack([oracle.dugo(*b, tl, "L")[0] for b in B])
    P = sy.cfg.P
    rep = lambda a: np.repeat(a[:, None], P, axis=1)  # noqa: E731
    (_, _, _), l00, l01, l10, lX, lY = _inputs(sy, pls, k, E)
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(ti, tl, tr)
        bulk1, st1 = ctx.tau2d(m00, m01, m10, X, Y)
        lay, st2 = ctx.tau2d_layered(rep(m00), rep(m01), rep(m10), X, Y)
        other, st3 = ctx.tau2d_layered(l00, l01, l10, lX, lY)
        bulk2, st4 = ctx.tau2d(m00, m01, m10, X, Y)
    assert (st1 == 0).all() and (st2 == 0).all() and (st3 == 0).all() and (st4 == 0).all()
    assert np.array_equal(bulk1, lay)
    assert np.array_equal(bulk1, bulk2)
    assert rel(other[0], bulk1[0]) > 1e-3  # the layered stack is a different system


def test_layered_tau2d_rejects_bad_arguments(gpu):
    sy, _ = LG.layered_system()
    ti, tl, tr = sy.tinv(sy.cfg.energies[0])
    n, P = sy.cfg.n, sy.cfg.P
    z = np.zeros((2, P, n, n), np.complex128)
    zx = np.zeros((2, n, n), np.complex128)
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(ti, tl, tr)
        with pytest.raises(ValueError):
            ctx.tau2d_layered(z[:, :P - 1], z, z, zx, zx)
        with pytest.raises(gpu.SkkrError):
            ctx.tau2d_layered(z[:0], z[:0], z[:0], zx[:0], zx[:0])
