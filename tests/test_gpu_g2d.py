"""tau2d and dugo with the reference's own arguments (skkr_tau2d_g2d,
skkr_dugo_g2d): g2d windows and padded leading dimensions as tau2d.f:2-3 and
dugo.f:9-10 declare them, packed on the GPU (packer.f, relmtrx.f).

The windows are filled from G(k) the way the reference driver that wrote the
golden vectors fills them (oracle/ref/refdrv.F, fillpl), per k-point, and
the results are checked against the reference's own per-k tau of the saved
k-points (tests/golden/*.npz, tau_save) and its decimation iteration counts.
The kappa-mu path (irel = 1) is pinned by the similarity it is: relmtrx is
C (a (x) 1_2) C^T with C the orthogonal Clebsch-Gordan matrix, so the
kappa-mu tau of a spin-independent system is relmtrx of its lm tau.

Tolerance: relative Frobenius error <= 1e-12 per k-point (FP64)."""
import importlib.util
import os

import numpy as np
import pytest

import oracle
from skkr_budapest_amd import systems as S

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TOL = 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def gpu():
    from skkr_budapest_amd import gpu as G
    return G


def gk(sy, k):
    """G_intra(k), G_inter(k) = sum_R G(R) exp(-i k.R)."""
    ph = np.exp(-1j * (sy.R @ k))
    return np.tensordot(ph, sy.G_intra, axes=(0, 0)), np.tensordot(ph, sy.G_inter, axes=(0, 0))


def fill_plane(g2d, plane, gi, gm, gn, lk, s):
    """One principal layer's window g2d(:, :, -minprc:minprc, 1:s, plane) from
    its intra block gi and its couplings to the PL before (gm^dagger) and
    after (gn), as refdrv.F's fillpl scatters them (0-based j + minprc)."""
    minprc = g2d.shape[3]
    blk = lambda m, a, b: m[a * lk:(a + 1) * lk, b * lk:(b + 1) * lk]
    for i in range(1, s + 1):
        for j in range(-s, -i + 1):
            jp = s + i + j
            g2d[:lk, :lk, j + minprc, i - 1, plane] = blk(gm, jp - 1, i - 1).conj().T
        for j in range(1 - i, s - i + 1):
            g2d[:lk, :lk, j + minprc, i - 1, plane] = blk(gi, i - 1, i + j - 1)
        for j in range(s - i + 1, s + 1):
            g2d[:lk, :lk, j + minprc, i - 1, plane] = blk(gn, i - 1, j + i - s - 1)


class Dims:
    """Padded dimensions of the reference's arrays (param.h's role): every
    one larger than the system needs."""

    def __init__(self, lk, s, P):
        self.lmmaxp, self.minprc, self.mprc1 = lk + 2, s + 1, P + 1
        self.lkmaxp, self.mintfc, self.mdimp = lk + 1, P * s + 3, lk * s + 5

    def g2d(self):
        return np.zeros((self.lmmaxp, self.lmmaxp, 2 * self.minprc + 1, self.minprc, self.mprc1 + 2),
                        np.complex128, order="F")

    def pad_t(self, t, nlay):
        """(nlay, lk, lk) [row, col] -> (lkmaxp, lkmaxp, mintfc) Fortran."""
        out = np.zeros((self.lkmaxp, self.lkmaxp, self.mintfc), np.complex128, order="F")
        lk = t.shape[-1]
        for l in range(nlay):
            out[:lk, :lk, l] = t[l]
        return out

    def square(self):
        return np.zeros((self.mdimp, self.mdimp), np.complex128, order="F")


def run_k(ctx, sy, E, k, d, irel=0, lmax=0, planes=None, tinv=None, fac=1.0, bulkgeo=True):
    """dugo('L'), dugo('R') and tau2d of one k-point through the
    reference-shaped entries; returns (tau (P*s, lk, lk), (iters L, iters R))."""
    cfg = sy.cfg
    lk, s, P = ctx.lk, cfg.s, cfg.P
    lkg = sy.G_intra.shape[-1] // s  # lm block size of G
    g2d = d.g2d()
    gi, go = gk(sy, k)
    for p in range(3):
        fill_plane(g2d, p, gi, go, go, lkg, s)
    gb = np.array(g2d[..., :3], order="F", copy=True)  # the bulk planes, kept for dugo
    if planes is not None:  # a layered stack: plane p for PL p
        for p, (pgi, pgm, pgn) in enumerate(planes, start=1):
            fill_plane(g2d, p, pgi, pgm, pgn, lkg, s)
    tint, tl, tr = tinv if tinv is not None else sy.tinv(E)
    ninprc = np.full(P + 2, s, np.int32)
    y, x = d.square(), d.square()
    rl, convl, itl = ctx.dugo_g2d(lmax, d.pad_t(tl, s)[:, :, :d.minprc], np.asfortranarray(gb[..., 0]),
                                  np.asfortranarray(gb[..., 1]), y, "L", True, irel, ninprc, fac)
    rr, convr, itr = ctx.dugo_g2d(lmax, d.pad_t(tr, s)[:, :, :d.minprc], np.asfortranarray(gb[..., 2]),
                                  np.asfortranarray(gb[..., 1]), x, "R", True, irel, ninprc, fac)
    assert rl == 0 and rr == 0 and convl and convr
    tau = np.full((d.lkmaxp, d.lkmaxp, d.mintfc), 7.0 + 0j, np.complex128, order="F")
    rc = ctx.tau2d_g2d(irel, 1.0, bulkgeo, lmax, d.pad_t(tint, P * s), g2d, x, y, tau, ninprc, fac)
    assert rc == 0
    assert (tau[:, :, P * s:] == 7.0).all() and (tau[lk:, :, :] == 7.0).all()  # padding untouched
    return np.moveaxis(tau[:lk, :lk, :P * s], 2, 0), (itl, itr)


def _golden_system(name):
    from test_oracle import GOLDEN
    return GOLDEN[name]


@pytest.mark.parametrize("name", ["cfg0_spin1", "cfg0_spin2", "herm16", "bdg18"])
def test_reference_shaped_kloop_matches_golden(gpu, name):
    g = np.load(os.path.join(GOLD, f"{name}.npz"))
    sy = _golden_system(name)
    cfg = sy.cfg
    E = complex(g["E"])
    d = Dims(cfg.lk, cfg.s, cfg.P)
    print(f"g2d entries: {name} lk={cfg.lk} s={cfg.s} P={cfg.P}", flush=True)
    with gpu.Context.for_system(sy) as ctx:
        for j, ik in enumerate(g["save_idx"]):
            tau, its = run_k(ctx, sy, E, g["k"][ik], d)
            assert tuple(g["iters"][ik]) == its
            assert rel(tau, g["tau_save"][j]) < TOL


def _layered_module():
    spec = importlib.util.spec_from_file_location("make_layered_golden",
                                                  os.path.join(GOLD, "make_layered_golden.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_reference_shaped_layered_tau2d_matches_golden(gpu):
    """bulkgeo = .false.: every PL reads its own g2d plane (tau2d.f:55-56,
    90-91), the leads the bulk planes."""
    LG = _layered_module()
    g = np.load(os.path.join(GOLD, "layered16.npz"))
    sy, pls = LG.layered_system()
    cfg = sy.cfg
    E = complex(g["E"])
    d = Dims(cfg.lk, cfg.s, cfg.P)
    print("g2d entries: layered16", flush=True)
    with gpu.Context.for_system(sy) as ctx:
        for j, ik in enumerate(g["save_idx"]):
            k = g["k"][ik]
            go_bulk = gk(sy, k)[1]
            own = [gk(p, k) for p in pls]
            planes = []
            for p in range(cfg.P):
                gm = own[p - 1][1] if p > 0 else go_bulk
                gn = own[p][1] if p < cfg.P - 1 else go_bulk
                planes.append((own[p][0], gm, gn))
            tau, its = run_k(ctx, sy, E, k, d, planes=planes, bulkgeo=False)
            assert tuple(g["iters"][ik]) == its
            assert rel(tau, g["tau_save"][j]) < TOL


def test_kappa_mu_blocks_are_the_similarity_of_the_lm_ones(gpu):
    """irel = 1 (relmtrx.f on every lm block of g2d, fac = 1): with t^-1 in
    the kappa-mu basis as relmtrx of the lm t^-1, tau_kappa-mu = relmtrx of
    the golden lm tau, and the decimation takes the same iterations."""
    g = np.load(os.path.join(GOLD, "cfg0_spin1.npz"))
    sy = _golden_system("cfg0_spin1")
    cfg = sy.cfg
    E = complex(g["E"])
    lmax = 2
    assert cfg.lk == (lmax + 1) ** 2
    lkr = 2 * cfg.lk
    d = Dims(lkr, cfg.s, cfg.P)
    d.lmmaxp = cfg.lk + 1       # the window holds lm blocks
    tint, tl, tr = sy.tinv(E)
    R = lambda t: np.array([oracle.relmtrx(b, 1.0, lmax) for b in t])
    tinv = (R(tint), R(tl), R(tr))
    print("g2d entries: kappa-mu cfg0_spin1", flush=True)
    with gpu.Context(lk=lkr, sites_per_pl=cfg.s, n_pl=cfg.P, errmax=cfg.errmax,
                     itermax=cfg.itermax) as ctx:
        for j, ik in enumerate(g["save_idx"][:2]):
            tau, its = run_k(ctx, sy, E, g["k"][ik], d, irel=1, lmax=lmax, tinv=tinv)
            assert tuple(g["iters"][ik]) == its
            assert rel(tau, R(g["tau_save"][j])) < TOL


def test_kappa_mu_fac_scales_the_structure_constants(gpu):
    """fac (common /relfac/) multiplies every lm block before relmtrx
    (relmtrx.f:27-31): the kappa-mu tau with fac = f and G equals the one
    with fac = 1 and f G, computed through the oracle's packer composition."""
    sy = _golden_system("cfg0_spin1")
    cfg = sy.cfg
    E = 0.5 + 0.05j
    lmax, f = 2, 0.9 + 0.05j
    lkr = 2 * cfg.lk
    d = Dims(lkr, cfg.s, cfg.P)
    d.lmmaxp = cfg.lk
    tint, tl, tr = sy.tinv(E)
    R = lambda t, fac=1.0: np.array([oracle.relmtrx(b, fac, lmax) for b in t])
    k = np.array([0.7, -1.9])
    with gpu.Context(lk=lkr, sites_per_pl=cfg.s, n_pl=cfg.P, errmax=cfg.errmax,
                     itermax=cfg.itermax) as ctx:
        tau, _ = run_k(ctx, sy, E, k, d, irel=1, lmax=lmax, tinv=(R(tint), R(tl), R(tr)), fac=f)
    # oracle: the lm packer blocks (the windows' blocks, negated), each lm
    # block through relmtrx with f (packer.f with irel = 1), dugo and tau2d
    gi, go = gk(sy, k)
    n9 = cfg.lk

    def relbig(m):
        nb = m.shape[0] // n9
        out = np.zeros((nb * lkr, nb * lkr), np.complex128)
        for a in range(nb):
            for b in range(nb):
                out[a * lkr:(a + 1) * lkr, b * lkr:(b + 1) * lkr] = oracle.relmtrx(
                    m[a * n9:(a + 1) * n9, b * n9:(b + 1) * n9], f, lmax)
        return out

    m00, m01, m10 = (relbig(m) for m in oracle.pl_blocks(gi, go))
    y, _, sl = oracle.dugo(m00, m01, m10, R(tl), "L", itermax=cfg.itermax, errmax=cfg.errmax)
    x, _, sr = oracle.dugo(m00, m01, m10, R(tr), "R", itermax=cfg.itermax, errmax=cfg.errmax)
    assert sl == 0 and sr == 0
    ref, st = oracle.tau2d(m00, m01, m10, R(tint), cfg.P, x, y)
    assert st == 0
    assert rel(tau, ref) < TOL


def test_reference_shaped_entries_reject_unequal_layers(gpu):
    sy = _golden_system("herm16")
    cfg = sy.cfg
    d = Dims(cfg.lk, cfg.s, cfg.P)
    ninprc = np.full(cfg.P + 2, cfg.s, np.int32)
    ninprc[2] = cfg.s - 1
    tau = np.zeros((d.lkmaxp, d.lkmaxp, d.mintfc), np.complex128, order="F")
    with gpu.Context.for_system(sy) as ctx:
        with pytest.raises(gpu.SkkrError):
            ctx.tau2d_g2d(0, 1.0, True, 3, d.pad_t(np.zeros((cfg.P * cfg.s, cfg.lk, cfg.lk)), cfg.P * cfg.s),
                          d.g2d(), d.square(), d.square(), tau, ninprc)
