"""The mixed-precision modes (INT8 digit products, split-BF16 products) on the
k-loop paths beside the Fourier producer: the layer structure constants
(skkr_set_structure_kkr, lm and kappa-mu bases) and the off-diagonal site-pair
sums (skkr_tau_bz_pairs), each against the FP64 oracle of the same inputs.

Tolerance: relative error <= 1e-8 of the BZ-integrated tau (and of the pair
sums), the accuracy the mixed modes are built for; the decimation iteration
counts stay within one of FP64's."""
import numpy as np
import pytest

import oracle
from skkr_budapest_amd import systems as S
from test_gpu_gstore import LMAX, _kkr_system

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-8
MODES = ["MIXED_INT8", "MIXED_BF16"]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def gpu():
    from skkr_budapest_amd import gpu as G
    return G


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("relm", [0, 1])
def test_mixed_kloop_on_layer_structure_constants(gpu, mode, relm):
    """Nb(110) stack of the reference's sample, lmax 2 (relm 0: lm basis, 9
    per site; relm 1: kappa-mu, 18 per site), 4 k-points, E = 0.5 + 0.05i."""
    G = gpu
    E = complex(0.5, 0.05)
    g, al, ti = _kkr_system(E)
    s, nl2 = g["s"], (LMAX + 1) ** 2
    P = 3 if relm == 0 else 2
    if relm:
        ti = np.array([oracle.relmtrx(t, 1.0, LMAX) for t in ti])
    tinv = np.concatenate([ti] * P)
    ks = np.array([[0.2, -0.1], [-0.45, 0.3], [0.05, 0.6], [0.33, 0.12]])
    ws = np.full(len(ks), 1.0 / len(ks))
    L = oracle.Lattice(g["a1"], g["a2"])
    ref, its = oracle.kkr_tau_bz(L, LMAX, s, g["nextra"], g["cvec"], al, E, g["eta"], tinv, ti, ti, ks, ws,
                                 rel=relm)
    lk = nl2 * (2 if relm else 1)
    print(f"{mode} k-loop on the layer producer, rel={relm}", flush=True)
    with G.Structure(LMAX, s, g["nextra"], g["eta"], g["a1"], g["a2"], g["cvec"]) as st:
        with G.Context(lk=lk, sites_per_pl=s, n_pl=P, precision=getattr(G, mode)) as ctx:
            ctx.set_tinv(tinv, ti, ti)
            if relm:
                ctx.set_structure_kkr(st, al, rel=relm)
            else:
                ctx.set_structure_kkr(st, al)
            r = ctx.tau_bz(E, ks, ws)
    assert (r["status"] == 0).all()
    assert np.abs(r["iters"] - its).max() <= 1
    e = rel(r["tau_bz"], ref)
    print("  tau rel. error", e, r["stats"]["dom_name"], flush=True)
    assert e < MIXED_TOL


@pytest.mark.parametrize("mode", MODES)
def test_mixed_pair_sums(gpu, mode):
    """Off-diagonal tau of site pairs with phase-weighted BZ sums (the cluster
    embedding input) in the mixed modes, against the dense-inverse oracle."""
    sy = S.make_system(S.small_config(lk=16, s=2, P=4, r2max=8, tmodel="herm", energies=(0.35 + 0.04j,)))
    E = sy.cfg.energies[0]
    k, w = S.uniform_kmesh(5)
    a1, a2 = (1.0, 0.0), (0.0, 1.0)
    ns = 4 * 2
    pairs = [(1, 1, 1, 0), (0, 1, 0, 0), (0, 5, 1, -1), (ns - 1, 0, 0, 2)]
    ref_ij, ref_ji = oracle.tau_pairs(sy, E, k, w, pairs, a1, a2)
    print(f"{mode} pair sums", flush=True)
    with gpu.Context.for_system(sy, precision=getattr(gpu, mode)) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        ctx.set_pairs(pairs, a1, a2)
        r = ctx.tau_bz_pairs(E, k, w)
    for q in range(len(pairs)):
        assert rel(r["tau_ij"][q], ref_ij[q]) < MIXED_TOL, q
        assert rel(r["tau_ji"][q], ref_ji[q]) < MIXED_TOL, q


def test_int8_tiles_give_the_same_kloop(gpu):
    """The INT8 k-loop on both work-group tiles (64 x 64, two per CU, and
    128 x 64, one per CU) of config 1 on 16 k-points: the same digit products
    and folds, so tau_BZ, the per-k traces and the decimation iteration counts
    are bitwise equal (the |alpha'|^2 partial sums of the stop rule are summed
    over different tile partitions; here they decide alike)."""
    cfg = S.CONFIGS[1]
    sy = S.make_system(cfg)
    E = cfg.energies[0]
    k, w = S.uniform_kmesh(cfg.mesh)
    idx = np.linspace(0, len(w) - 1, 16).astype(int)
    ks, ws = k[idx], np.full(len(idx), 1.0 / len(idx))
    out = {}
    for tile in (64, 128):
        print(f"INT8 k-loop, tile {tile}", flush=True)
        with gpu.Context.for_system(sy, precision=gpu.MIXED_INT8) as ctx:
            ctx.set_option("int8_tile", tile)
            ctx.set_tinv(*sy.tinv(E))
            out[tile] = ctx.tau_bz(E, ks, ws)
    a, b = out[64], out[128]
    assert (a["status"] == 0).all() and (b["status"] == 0).all()
    np.testing.assert_array_equal(a["iters"], b["iters"])
    np.testing.assert_array_equal(a["tau_bz"], b["tau_bz"])
    np.testing.assert_array_equal(a["trk"], b["trk"])


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("lk,s,P", [(9, 1, 3), (18, 2, 4), (24, 3, 3), (20, 5, 2), (26, 5, 2)])
def test_mixed_kloop_ragged_shapes(gpu, mode, lk, s, P):
    """The mixed k-loop on PL widths n = 9, 36, 72, 100, 130: partial tiles,
    partial K-chunks, decimation products whose tiles straddle the ε / α'
    halves (rows n - 1 | n), and n > 128 (the 2 x 2 block inversion); 9
    k-points with an uneven last batch of 4 (batch 5).  Against the oracle."""
    sy = S.make_system(S.small_config(lk=lk, s=s, P=P, r2max=4, tmodel="herm",
                                      energies=(0.45 + 0.05j,), cfg_id=95))
    E = sy.cfg.energies[0]
    rng = np.random.default_rng(lk * 100 + s)
    k = rng.uniform(-np.pi, np.pi, (9, 2))
    w = np.full(9, 1.0 / 9)
    ref = oracle.tau_bz(sy, E, k, w, nworkers=8)
    print(f"{mode} ragged k-loop n={lk * s}", flush=True)
    with gpu.Context.for_system(sy, precision=getattr(gpu, mode), batch=5) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        r = ctx.tau_bz(E, k, w)
    assert (r["status"] == 0).all()
    assert np.abs(r["iters"] - ref["iters"]).max() <= 1
    e = rel(r["tau_bz"], ref["tau_bz"])
    print("  tau rel. error", e, flush=True)
    assert e < MIXED_TOL


@pytest.mark.parametrize("mode", MODES)
def test_mixed_physical_tau(gpu, mode):
    """skkr_tau_bz_phys in the mixed modes: the physical BZ sum and the per-k
    per-l traces of phystau (cpacoord.f:294-311) against the oracle's
    phystau of the FP64 oracle k-loop, within 1e-8."""
    sy = S.make_system(S.small_config(lk=8, s=2, P=3, r2max=4, tmodel="herm", energies=(0.35 + 0.04j,)))
    E = sy.cfg.energies[0]
    k, w = S.uniform_kmesh(4)
    nl, lk, ns = 3, 8, 6
    rng = np.random.default_rng(8)
    lidx = np.sort(np.concatenate([np.arange(nl), rng.integers(0, nl, size=lk - nl)]))
    a = (0.2 * (rng.normal(size=(ns, nl)) + 1j * rng.normal(size=(ns, nl))))[:, lidx]
    ref = oracle.tau_bz(sy, E, k, w, nworkers=8, want_tau_k=True)
    tint = sy.tinv(E)[0]
    print(f"{mode} physical tau", flush=True)
    with gpu.Context.for_system(sy, precision=getattr(gpu, mode)) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        ctx.set_screening(a, lidx, nl)
        r = ctx.tau_bz_phys(E, k, w)
    phys_ref = np.stack([oracle.phystau(ref["tau_bz"][i], tint[i], tint[i], a[i]) for i in range(ns)])
    assert rel(r["tau_bz_phys"], phys_ref) < MIXED_TOL
    tkd_ref = np.array([[oracle.phys_traces(ref["tau_k"][ik, i], tint[i], a[i], lidx, nl)
                         for i in range(ns)] for ik in range(len(w))])
    assert rel(r["taukdiag"], tkd_ref) < MIXED_TOL
