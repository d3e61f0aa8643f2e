"""Parity of the HIP path (through the C-ABI) with the CPU oracle and the
reference's golden vectors.  FP64 tolerance: relative Frobenius error
<= 1e-11 for tau (round-off of ~1e3 dependent complex operations per k),
identical decimation iteration counts."""
import os

import numpy as np
import pytest

import oracle
from skkr_budapest_amd import systems as S

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TOL = 1e-11
# config 1 (n=128, 20 PLs, Im E=0.02): the Schur complements and decimation
# blocks have condition numbers ~2e3 and each k chains ~60 dependent
# inversions, so two FP64 implementations with different pivot orders
# (column pivoting in the oracle, block-restricted row pivoting on the GPU)
# differ by ~1e-11 relative
TOL_CFG1 = 5e-11


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def gpu():
    from skkr_budapest_amd import gpu as G
    return G


def small_sys(lk=4, s=2, P=3, tmodel="herm", E=0.3 + 0.1j, r2max=2):
    cfg = S.small_config(lk=lk, s=s, P=P, r2max=r2max, tmodel=tmodel, energies=(E,))
    return S.make_system(cfg)


@pytest.mark.parametrize("n", [9, 18, 36, 64, 65, 72, 96, 100, 128])
def test_zgemm_matches_numpy(gpu, n):
    rng = np.random.default_rng(n)
    a = rng.normal(size=(5, n, n)) + 1j * rng.normal(size=(5, n, n))
    b = rng.normal(size=(5, n, n)) + 1j * rng.normal(size=(5, n, n))
    with gpu.Context(lk=n, sites_per_pl=1, n_pl=1) as ctx:
        c = ctx.zgemm(a, b)
    ref = a @ b
    assert rel(c, ref) < 1e-14


def test_zgemm_every_tile_on_ragged_batches(gpu):
    """Each work-group tile forced on a ragged shape (n = 72 and 200) and a
    batch of 11, so that the XCD renumbering (8 products at a time, the rest
    plain) meets several tiles per product.  All tiles of one product form
    run the same K order and MFMA instruction: their results are bitwise
    equal."""
    rng = np.random.default_rng(96)
    for n in (72, 200):
        a = rng.normal(size=(11, n, n)) + 1j * rng.normal(size=(11, n, n))
        b = rng.normal(size=(11, n, n)) + 1j * rng.normal(size=(11, n, n))
        ref = a @ b
        first = {}
        for tile in (32, 64, 96, 128):
            with gpu.Context(lk=n, sites_per_pl=1, n_pl=1) as ctx:
                ctx.set_option("gemm_tile", tile)
                c = ctx.zgemm(a, b)
                ctx.set_option("gemm_3m", 1)
                c3 = ctx.zgemm(a, b)
            assert rel(c, ref) < 1e-14
            assert rel(c3, ref) < 1e-14
            first.setdefault("plain", c)
            first.setdefault("3m", c3)
            np.testing.assert_array_equal(c, first["plain"])
            np.testing.assert_array_equal(c3, first["3m"])


@pytest.mark.parametrize("n", [64, 128, 200])
def test_zgemm_three_product_matches_numpy(gpu, n):
    """The Gauss three-product complex GEMM (option gemm_3m, 128 x 64 tiles):
    normwise FP64 accuracy, ragged edges included (n = 200)."""
    rng = np.random.default_rng(300 + n)
    a = rng.normal(size=(3, n, n)) + 1j * rng.normal(size=(3, n, n))
    b = rng.normal(size=(3, n, n)) + 1j * rng.normal(size=(3, n, n))
    with gpu.Context(lk=n, sites_per_pl=1, n_pl=1) as ctx:
        ctx.set_option("gemm_3m", 1)
        c3 = ctx.zgemm(a, b)
        ctx.set_option("gemm_3m", 0)
        c4 = ctx.zgemm(a, b)
    ref = a @ b
    assert rel(c3, ref) < 1e-14
    assert rel(c4, ref) < 1e-14


@pytest.mark.parametrize("kernel", ["blocked", "rowwise"])
@pytest.mark.parametrize("n", [1, 9, 18, 36, 48, 64, 65, 72, 80, 90, 96, 100, 112, 128])
def test_gjinv_matches_numpy(gpu, n, kernel, monkeypatch):
    monkeypatch.setenv("SKKR_INV_KERNEL", kernel)
    rng = np.random.default_rng(100 + n)
    a = rng.normal(size=(7, n, n)) + 1j * rng.normal(size=(7, n, n))
    a[1, 0, 0] = 1e-14  # forces pivoting
    with gpu.Context(lk=n, sites_per_pl=1, n_pl=1) as ctx:
        inv, st = ctx.gjinv(a)
        assert (st == 0).all()
        res = np.abs(inv @ a - np.eye(n)).max()
        assert res < 1e-9
        z, st = ctx.gjinv(np.zeros((2, n, n)))
        assert (st == 4).all()


@pytest.mark.parametrize("n", [9, 36, 64, 72, 96, 112, 128])
def test_gjinv_without_diagonal_pivoting_on_accretive_matrices(gpu, n):
    """inv_kernel 2 (no row exchanges inside the 16x16 diagonal blocks) on
    matrices of the layer problem's class: E - H with Im E > 0, whose
    anti-Hermitian part is positive definite, so every leading block is
    non-singular.  The tolerance is looser than with pivoting."""
    rng = np.random.default_rng(500 + n)
    a = rng.normal(size=(6, n, n)) + 1j * rng.normal(size=(6, n, n))
    h = (a + np.conj(np.swapaxes(a, 1, 2))) / 2
    m = (0.3 + 0.02j) * np.eye(n) - h
    with gpu.Context(lk=n, sites_per_pl=1, n_pl=1) as ctx:
        ctx.set_option("inv_kernel", 2)
        inv, st = ctx.gjinv(m)
        assert (st == 0).all()
        ref = np.linalg.inv(m)
        assert rel(inv, ref) < 1e-10
        z, st = ctx.gjinv(np.zeros((2, n, n)))
        assert (st == 4).all()


def test_config1_sample_without_diagonal_pivoting(gpu):
    """Config 1 with inv_kernel 2 in the decimation and sweeps."""
    cfg = S.CONFIGS[1]
    sy = S.make_system(cfg)
    E = cfg.energies[0]
    k, w = S.uniform_kmesh(cfg.mesh)
    idx = np.array([0, 1300, 2048, 3500])
    ks, ws = k[idx], w[idx]
    ref = oracle.tau_bz(sy, E, ks, ws, nworkers=8)
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_option("inv_kernel", 2)
        ctx.set_tinv(*sy.tinv(E))
        r = ctx.tau_bz(E, ks, ws)
    np.testing.assert_array_equal(r["iters"], ref["iters"])
    assert rel(r["tau_bz"], ref["tau_bz"]) < TOL_CFG1


def test_pl_blocks_match_oracle(gpu):
    sy = S.make_system(S.small_config(lk=6, s=3, P=2, r2max=8))
    rng = np.random.default_rng(3)
    k = rng.uniform(-np.pi, np.pi, size=(40, 2))
    with gpu.Context.for_system(sy) as ctx:
        m00, m01, m10 = ctx.pl_blocks(k)
    for i in range(0, 40, 7):
        gi = oracle.fourier(sy.R, sy.G_intra, k[i])
        go = oracle.fourier(sy.R, sy.G_inter, k[i])
        r00, r01, r10 = oracle.pl_blocks(gi, go)
        for a, b in ((m00[i], r00), (m01[i], r01), (m10[i], r10)):
            assert np.abs(a - b).max() < 1e-14 * max(1.0, np.abs(b).max())


def _blocks(sy, k):
    gi = oracle.fourier(sy.R, sy.G_intra, k)
    go = oracle.fourier(sy.R, sy.G_inter, k)
    return oracle.pl_blocks(gi, go)


def test_surfgfd_and_dugo_match_oracle(gpu):
    sy = small_sys(lk=5, s=2, P=2)
    E = sy.cfg.energies[0]
    ti, tl, tr = sy.tinv(E)
    ks = np.array([[0.1, 0.2], [-1.0, 2.5], [3.0, -0.4]])
    B = [_blocks(sy, k) for k in ks]
    m00 = np.stack([b[0] for b in B]); m01 = np.stack([b[1] for b in B])
    m10 = np.stack([b[2] for b in B])
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(ti, tl, tr)
        # plain surfgfd on m00 + t_left
        m0 = m00.copy()
        for i in range(3):
            full = m00[i].copy()
            for a in range(sy.s):
                full[a * sy.lk:(a + 1) * sy.lk, a * sy.lk:(a + 1) * sy.lk] += tl[a]
            m0[i] = full
        x, it, st = ctx.surfgfd(m10, m0, m01)
        for i in range(3):
            xr, itr, str_ = oracle.surfgfd(m10[i], m0[i], m01[i])
            assert it[i] == itr and st[i] == 0
            assert rel(x[i], xr) < TOL
        for side in ("L", "R"):
            xy, it, st = ctx.dugo(side, m00, m01, m10)
            for i in range(3):
                ref, itr, _ = oracle.dugo(m00[i], m01[i], m10[i], tl if side == "L" else tr, side)
                assert it[i] == itr and st[i] == 0
                assert rel(xy[i], ref) < TOL


def test_tau2d_matches_oracle(gpu):
    sy = small_sys(lk=4, s=3, P=5)
    E = sy.cfg.energies[0]
    ti, tl, tr = sy.tinv(E)
    ks = np.array([[0.3, -0.2], [1.7, 0.9]])
    B = [_blocks(sy, k) for k in ks]
    m00 = np.stack([b[0] for b in B]); m01 = np.stack([b[1] for b in B])
    m10 = np.stack([b[2] for b in B])
    X = np.stack([oracle.dugo(*B[i], tr, "R")[0] for i in range(2)])
    Y = np.stack([oracle.dugo(*B[i], tl, "L")[0] for i in range(2)])
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(ti, tl, tr)
        tau, st = ctx.tau2d(m00, m01, m10, X, Y, w=0.5)
    assert (st == 0).all()
    for i in range(2):
        ref, _ = oracle.tau2d(m00[i], m01[i], m10[i], ti, sy.P, X[i], Y[i], w=0.5)
        assert rel(tau[i], ref) < TOL


def _golden_system(name):
    from test_oracle import GOLDEN
    return GOLDEN[name]


@pytest.mark.parametrize("name", ["cfg0_spin1", "cfg0_spin2", "bdg18", "herm16"])
@pytest.mark.parametrize("batch", [0, 5])
def test_tau_bz_matches_reference_golden(gpu, name, batch):
    g = np.load(os.path.join(GOLD, f"{name}.npz"))
    sy = _golden_system(name)
    E = complex(g["E"])
    with gpu.Context.for_system(sy, batch=batch) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        r = ctx.tau_bz(E, g["k"], g["w"])
    np.testing.assert_array_equal(r["iters"], g["iters"])
    assert (r["status"] == 0).all()
    assert rel(r["tau_bz"], g["tau_bz"]) < TOL
    assert np.abs(r["trk"] - g["trk"]).max() < TOL * np.abs(g["trk"]).max()


def test_config0_full_mesh_matches_oracle(gpu):
    cfg = S.CONFIGS[0]
    sy = S.make_system(cfg)
    E = cfg.energies[0]
    k, w = S.uniform_kmesh(cfg.mesh)
    ref = oracle.tau_bz(sy, E, k, w, nworkers=8)
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        r = ctx.tau_bz(E, k, w)
    np.testing.assert_array_equal(r["iters"], ref["iters"])
    assert rel(r["tau_bz"], ref["tau_bz"]) < TOL
    assert np.abs(r["trk"] - ref["trk"]).max() < TOL * np.abs(ref["trk"]).max()


def test_config1_sample_matches_oracle(gpu):
    """Config 1 (n=128, 20 PLs, 25 R) on 8 k-points of its 64x64 mesh."""
    cfg = S.CONFIGS[1]
    sy = S.make_system(cfg)
    E = cfg.energies[0]
    k, w = S.uniform_kmesh(cfg.mesh)
    idx = np.array([0, 517, 1300, 2047, 2048, 3001, 3500, 4095])
    ks, ws = k[idx], w[idx]
    ref = oracle.tau_bz(sy, E, ks, ws, nworkers=8)
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        r = ctx.tau_bz(E, ks, ws)
    np.testing.assert_array_equal(r["iters"], ref["iters"])
    assert rel(r["tau_bz"], ref["tau_bz"]) < TOL_CFG1
    assert np.abs(r["trk"] - ref["trk"]).max() < TOL_CFG1 * np.abs(ref["trk"]).max()


def test_config1_sample_with_three_product_gemm(gpu):
    """Config 1 with the three-product GEMM in the decimation and sweeps."""
    cfg = S.CONFIGS[1]
    sy = S.make_system(cfg)
    E = cfg.energies[0]
    k, w = S.uniform_kmesh(cfg.mesh)
    idx = np.array([0, 1300, 2048, 3500])
    ks, ws = k[idx], w[idx]
    ref = oracle.tau_bz(sy, E, ks, ws, nworkers=8)
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_option("gemm_3m", 1)
        ctx.set_tinv(*sy.tinv(E))
        r = ctx.tau_bz(E, ks, ws)
    np.testing.assert_array_equal(r["iters"], ref["iters"])
    assert rel(r["tau_bz"], ref["tau_bz"]) < TOL_CFG1


def test_config1_full_mesh_properties(gpu):
    """Full config-1 energy (4096 k): per-k results do not depend on how the k
    are batched (bitwise), and the BZ sum is linear in the weights."""
    cfg = S.CONFIGS[1]
    sy = S.make_system(cfg)
    E = cfg.energies[0]
    k, w = S.uniform_kmesh(cfg.mesh)
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        r = ctx.tau_bz(E, k, w)
    assert (r["status"] == 0).all()
    with gpu.Context.for_system(sy, batch=97) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        sel = np.arange(0, 4096, 13)
        r2 = ctx.tau_bz(E, k[sel], w[sel])
    np.testing.assert_array_equal(r2["trk"], r["trk"][sel])
    np.testing.assert_array_equal(r2["iters"], r["iters"][sel])
    tr_bz = np.einsum("sll->s", r["tau_bz"])
    assert rel(tr_bz, (w[:, None] * r["trk"]).sum(0)) < 1e-13
    # DOS of each site is positive: -Im tr tau / pi > 0 (resolvent, Im E > 0)
    assert (-tr_bz.imag > 0).all()


def test_itermax_exceeded_is_reported(gpu):
    sy = small_sys(lk=3, s=2, P=2, E=0.2 + 0.001j)
    E = sy.cfg.energies[0]
    k, w = S.uniform_kmesh(2)
    with gpu.Context.for_system(sy, itermax=1) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        r = ctx.tau_bz(E, k, w, raise_on_error=False)
    assert r["rc"] == 6
    assert (r["status"] == 6).all()
    ref = oracle.tau_bz(sy, E, k, w, itermax=1)
    assert ref["rc"] == 3
    np.testing.assert_array_equal(r["iters"], ref["iters"])


@pytest.mark.parametrize("n", [144, 200, 256])
def test_block_inverse_beyond_128(gpu, n):
    """n > 128: 2x2 block elimination around the leading 128 x 128 block, for
    matrices with a positive-definite anti-Hermitian part (E - H, Im E > 0)."""
    rng = np.random.default_rng(n)
    h = rng.normal(size=(4, n, n)) + 1j * rng.normal(size=(4, n, n))
    h = (h + np.conj(np.swapaxes(h, 1, 2))) / 2
    a = (0.3 + 0.05j) * np.eye(n) - h
    with gpu.Context(lk=n, sites_per_pl=1, n_pl=1) as ctx:
        inv, st = ctx.gjinv(a)
    assert (st == 0).all()
    ref = np.linalg.inv(a)
    assert rel(inv, ref) < 1e-11


def test_config3_sample_matches_oracle(gpu):
    """Config 3 (BdG, lk=64, n=256, 12 PLs) on 4 k-points at one contour energy."""
    from skkr_budapest_amd import contour
    cfg = S.CONFIGS[3]
    sy = S.make_system(cfg)
    z, _ = contour.config_contour()
    E = z[13]
    k, w = S.uniform_kmesh(16)
    idx = np.array([0, 77, 130, 255])
    ks, ws = k[idx], w[idx]
    ref = oracle.tau_bz(sy, E, ks, ws, nworkers=4)
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        r = ctx.tau_bz(E, ks, ws)
    np.testing.assert_array_equal(r["iters"], ref["iters"])
    assert rel(r["tau_bz"], ref["tau_bz"]) < TOL_CFG1


def test_graph_replay_and_tiles_are_bitwise_identical(gpu):
    """HIP-graph replays of the decimation rounds and the sweeps, forced GEMM
    tiles and the autotuned choice all give the same FP64 bits (and decimation
    iteration counts) as the eager default."""
    sy = small_sys(lk=16, s=4, P=6, r2max=4, E=0.4 + 0.05j)
    E = sy.cfg.energies[0]
    k, w = S.uniform_kmesh(6)
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        ctx.set_option("graphs", 0)  # graphs are the default: the eager launches first
        base = ctx.tau_bz(E, k, w)
        ctx.set_option("graphs", 1)
        for rep in range(3):
            r = ctx.tau_bz(E, k, w)
            np.testing.assert_array_equal(r["tau_bz"], base["tau_bz"])
            np.testing.assert_array_equal(r["trk"], base["trk"])
            np.testing.assert_array_equal(r["iters"], base["iters"])
            assert r["stats"]["n_launches"] < base["stats"]["n_launches"]
            # eager work on another batch size between replays
            ctx.set_option("graphs", 0)
            ctx.tau_bz(E, k[:7], w[:7])
            ctx.set_option("graphs", 1)
        ctx.set_option("graphs", 0)
        for tile in (32, 64, 128):
            ctx.set_option("gemm_tile", tile)
            r = ctx.tau_bz(E, k, w)
            np.testing.assert_array_equal(r["tau_bz"], base["tau_bz"])
        ctx.set_option("gemm_tile", 0)
        t = ctx.autotune()
        assert t["tile_nn"] in (32, 64, 128) and t["tile_2nn"] in (32, 64, 128)
        r = ctx.tau_bz(E, k, w)
        np.testing.assert_array_equal(r["tau_bz"], base["tau_bz"])


def test_graph_replay_mixed_int8_bitwise(gpu):
    """The INT8 mixed mode replays its decimation rounds and sweeps from graphs
    with the bits of its eager launches, at two energies in turn (different
    round lengths)."""
    sy = small_sys(lk=16, s=4, P=6, r2max=4, E=0.4 + 0.05j)
    k, w = S.uniform_kmesh(6)
    with gpu.Context.for_system(sy, precision=gpu.MIXED_INT8) as ctx:
        for E in (0.4 + 0.05j, 0.1 + 0.3j, 0.4 + 0.05j):
            ctx.set_tinv(*sy.tinv(E))
            ctx.set_option("graphs", 0)
            base = ctx.tau_bz(E, k, w)
            ctx.set_option("graphs", 1)
            for rep in range(2):
                r = ctx.tau_bz(E, k, w)
                np.testing.assert_array_equal(r["tau_bz"], base["tau_bz"])
                np.testing.assert_array_equal(r["iters"], base["iters"])


def test_smaller_call_after_a_larger_batch(gpu):
    """A call with fewer k-points than the batch an earlier call allocated
    (the adaptive mesh does this level by level) gives the same bits as a
    fresh context."""
    sy = small_sys(lk=8, s=2, P=4, r2max=4, E=0.35 + 0.03j)
    E = sy.cfg.energies[0]
    k, w = S.uniform_kmesh(12)                      # 144 k
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        ctx.tau_bz(E, k, w)                         # allocates a batch of 144
        small = ctx.tau_bz(E, k[5:22], w[5:22])     # 17 k
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        fresh = ctx.tau_bz(E, k[5:22], w[5:22])
    np.testing.assert_array_equal(small["trk"], fresh["trk"])
    np.testing.assert_array_equal(small["iters"], fresh["iters"])
    np.testing.assert_array_equal(small["tau_bz"], fresh["tau_bz"])


def _screening(sy, nl, seed):
    rng = np.random.default_rng(seed)
    lk, ns = sy.cfg.lk, sy.cfg.P * sy.cfg.s
    lidx = np.sort(rng.integers(0, nl, size=lk))
    lidx[:nl] = np.arange(nl)
    lidx = np.sort(lidx)
    alpha_l = 0.2 * (rng.normal(size=(ns, nl)) + 1j * rng.normal(size=(ns, nl)))
    return alpha_l[:, lidx], lidx      # per-index parameters, like lml / ldex


@pytest.mark.parametrize("lk,s,P,tmodel", [(8, 2, 3, "herm"), (32, 1, 2, "herm"),
                                           (36, 1, 2, "bdg")])
def test_physical_tau_matches_oracle(gpu, lk, s, P, tmodel):
    """skkr_tau_bz_phys: phystau of the BZ sum and the per-k per-l traces of
    phystau(tau(k)) (cpacoord.f:294-311) against the oracle's phystau, whose
    output is pinned to the reference's phystau in test_oracle.py."""
    sy = small_sys(lk=lk, s=s, P=P, r2max=4, tmodel=tmodel, E=0.35 + 0.04j)
    E = sy.cfg.energies[0]
    k, w = S.uniform_kmesh(4)
    nl = 3
    a, lidx = _screening(sy, nl, seed=lk)
    ref = oracle.tau_bz(sy, E, k, w, nworkers=8, want_tau_k=True)
    tint = sy.tinv(E)[0]
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        ctx.set_screening(a, lidx, nl)
        r = ctx.tau_bz_phys(E, k, w)
        plain = ctx.tau_bz(E, k, w)
    np.testing.assert_array_equal(r["tau_bz"], plain["tau_bz"])
    ns = P * s
    phys_ref = np.stack([oracle.phystau(ref["tau_bz"][i], tint[i], tint[i], a[i]) for i in range(ns)])
    assert rel(r["tau_bz_phys"], phys_ref) < TOL
    tkd_ref = np.array([[oracle.phys_traces(ref["tau_k"][ik, i], tint[i], a[i], lidx, nl)
                         for i in range(ns)] for ik in range(len(w))])
    assert r["taukdiag"].shape == (len(w), ns, nl)
    assert rel(r["taukdiag"], tkd_ref) < TOL
    # the BZ-weighted per-k traces are the traces of the physical BZ sum
    tr_bz = np.array([[np.trace(phys_ref[i][np.ix_(lidx == l, lidx == l)]) for l in range(nl)]
                      for i in range(ns)])
    assert rel(np.einsum("k,kil->il", w, r["taukdiag"]), tr_bz) < TOL


@pytest.mark.parametrize("lk,s", [(18, 4), (24, 4)])
def test_kloop_in_the_96_tile_range_matches_oracle(gpu, lk, s):
    """n = 72 and 96: the products run in 96 x 96 work-group tiles (9 waves,
    whose norm reduction folds the threads above 512 first) and the
    inversions in the 80- and 96-row kernels; tau and the decimation
    iteration counts equal the oracle's."""
    sy = small_sys(lk=lk, s=s, P=3, r2max=4, E=0.35 + 0.04j)
    E = sy.cfg.energies[0]
    k, w = S.uniform_kmesh(4)
    ref = oracle.tau_bz(sy, E, k, w, nworkers=8)
    print("k-loop n =", lk * s, "P = 3, 16 k", flush=True)
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        r = ctx.tau_bz(E, k, w)
    np.testing.assert_array_equal(r["iters"], ref["iters"])
    assert rel(r["tau_bz"], ref["tau_bz"]) < TOL


def test_physical_tau_adds_up_over_calls(gpu):
    sy = small_sys(lk=8, s=2, P=3, r2max=4, E=0.35 + 0.04j)
    E = sy.cfg.energies[0]
    k, w = S.uniform_kmesh(4)
    a, lidx = _screening(sy, 3, seed=1)
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        ctx.set_screening(a, lidx, 3)
        whole = ctx.tau_bz_phys(E, k, w)
        p1 = ctx.tau_bz_phys(E, k[:5], w[:5])
        p2 = ctx.tau_bz_phys(E, k[5:], w[5:])
    assert rel(p1["tau_bz_phys"] + p2["tau_bz_phys"], whole["tau_bz_phys"]) < 1e-13
    np.testing.assert_array_equal(np.concatenate([p1["taukdiag"], p2["taukdiag"]]),
                                  whole["taukdiag"])


def test_empty_and_single_k_calls(gpu):
    """tau_bz accumulates (+=): a call with no k-points returns OK and leaves
    the sum as it was; one k-point alone matches the oracle; the same k twice
    in one call gives twice the sum of one (the rank-partial-sum contract
    of the reference's k-loop, cpacoord.f:147-149,326)."""
    sy = small_sys(lk=8, s=2, P=3, r2max=4, E=0.35 + 0.04j)
    E = sy.cfg.energies[0]
    k1, w1 = np.array([[0.21, -0.37]]), np.array([1.0])
    ref = oracle.tau_bz(sy, E, k1, w1)
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        one = ctx.tau_bz(E, k1, w1)
        assert one["rc"] == 0 and (one["status"] == 0).all()
        assert rel(one["tau_bz"], ref["tau_bz"]) < TOL
        acc = one["tau_bz"].copy()
        empty = ctx.tau_bz(E, np.zeros((0, 2)), np.zeros(0), tau_bz=acc)
        assert empty["rc"] == 0 and empty["trk"].shape[0] == 0
        np.testing.assert_array_equal(acc, one["tau_bz"])
        two = ctx.tau_bz(E, np.concatenate([k1, k1]), np.concatenate([w1, w1]))
    assert rel(two["tau_bz"], 2 * one["tau_bz"]) < 1e-14
    np.testing.assert_array_equal(two["iters"][0], two["iters"][1])


def test_set_option_rejects_unknown(gpu):
    with gpu.Context(lk=4, sites_per_pl=1, n_pl=1) as ctx:
        with pytest.raises(gpu.SkkrError):
            ctx.set_option("no_such_option", 1)
        with pytest.raises(gpu.SkkrError):
            ctx.set_option("gemm_tile", 48)


def test_energy_loop_matches_oracle_per_energy(gpu):
    """Contour loop with per-energy meshes (set_tinv per energy), checked per
    energy against the oracle, and the contour integral of tr G."""
    from skkr_budapest_amd import contour, energy_loop, observables as O
    sy = small_sys(lk=8, s=2, P=4, r2max=4)
    z, wz = contour.semicircle(6, -0.8, 0.6, 1.0)
    sides = energy_loop.geometric_mesh_sides(z.imag, 16, 64)
    refs = []
    for E, side in zip(z, sides):
        k, w = S.uniform_kmesh(side)
        refs.append(oracle.tau_bz(sy, E, k, w, nworkers=4)["tau_bz"])
    with gpu.Context.for_system(sy) as ctx:
        recs, total = energy_loop.run_contour(ctx, sy, z, wz, sides, reference=refs)
    for r in recs:
        assert r["tau_rel_err"] < TOL
        assert r["dos"] > 0
    ref_total = O.integrated_g([O.tr_g(t) for t in refs], wz)
    assert abs(total - ref_total) < TOL * abs(ref_total)
