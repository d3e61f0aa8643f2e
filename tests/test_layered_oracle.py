"""The oracle's tau2d on a layered stack (bulkgeo false) against the
reference's own tau2d (skkr_dbogo_bsf/main/tau2d.f with one g2d plane per
principal layer, tests/golden/layered16.npz)."""
import importlib.util
import os

import numpy as np

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _golden_module():
    spec = importlib.util.spec_from_file_location("make_layered_golden",
                                                  os.path.join(GOLD, "make_layered_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


LG = _golden_module()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def layered_blocks(sy, pls, k):
    """Lead blocks of the bulk structure and the packer blocks of every
    interface PL: M00_p and M01_p from PL p's own plane, M10_p the coupling
    of PL p-1 to p read back (the leads' bulk coupling at the two ends)."""
    lead = oracle.pl_blocks(*LG.gk(sy, k))
    go_bulk = LG.gk(sy, k)[1]
    P = len(pls)
    own = [LG.gk(p, k) for p in pls]
    m00 = np.stack([oracle.pl_blocks(gi, go)[0] for gi, go in own])
    m01 = np.stack([oracle.pl_blocks(own[p][0], own[p][1] if p < P - 1 else go_bulk)[1]
                    for p in range(P)])
    m10 = np.stack([oracle.pl_blocks(own[p - 1][0], own[p - 1][1])[2] if p > 0
                    else oracle.pl_blocks(own[0][0], go_bulk)[2] for p in range(P)])
    return lead, m00, m01, m10


def layered_tau(sy, pls, k, E, shift_m10=False):
    ti, tl, tr = sy.tinv(E)
    lead, m00, m01, m10 = layered_blocks(sy, pls, k)
    if shift_m10:  # a planted mistake: each PL reads the next plane's M10
        m10 = np.roll(m10, -1, axis=0)
    y, ity, _ = oracle.dugo(*lead, tl, "L")
    x, itx, _ = oracle.dugo(*lead, tr, "R")
    tau, st = oracle.tau2d_layered(m00, m01, m10, ti, x, y)
    assert st == 0
    return tau, (ity, itx)


def test_layered_system_is_the_golden_one():
    g = np.load(os.path.join(GOLD, "layered16.npz"))
    sy, pls = LG.layered_system()
    assert list(g["fingerprint"]) == [sy.fingerprint()] + [p.fingerprint() for p in pls]


def test_oracle_layered_tau2d_matches_reference():
    g = np.load(os.path.join(GOLD, "layered16.npz"))
    sy, pls = LG.layered_system()
    E = complex(g["E"])
    k, w = g["k"], g["w"]
    bz = np.zeros_like(g["tau_bz"])
    for ik in range(len(w)):
        tau, its = layered_tau(sy, pls, k[ik], E)
        assert tuple(g["iters"][ik]) == its
        bz += w[ik] * tau
        if ik in g["save_idx"]:
            j = list(g["save_idx"]).index(ik)
            assert rel(tau, g["tau_save"][j]) < 1e-12
    assert rel(bz, g["tau_bz"]) < 1e-12


def test_planted_plane_shift_is_caught():
    """Reading PL p's coupling from the wrong plane changes tau far beyond the
    bound above, so the golden pins the per-PL indexing."""
    g = np.load(os.path.join(GOLD, "layered16.npz"))
    sy, pls = LG.layered_system()
    j, ik = 1, int(g["save_idx"][1])
    tau, _ = layered_tau(sy, pls, g["k"][ik], complex(g["E"]), shift_m10=True)
    assert rel(tau, g["tau_save"][j]) > 1e-3


def test_layered_with_equal_planes_is_bulk_tau2d():
    """One plane repeated for every PL reproduces the bulkgeo tau2d exactly."""
    sy, _ = LG.layered_system()
    E = sy.cfg.energies[0]
    ti, tl, tr = sy.tinv(E)
    k = np.array([0.7, -1.1])
    m00, m01, m10 = oracle.pl_blocks(*LG.gk(sy, k))
    y = oracle.dugo(m00, m01, m10, tl, "L")[0]
    x = oracle.dugo(m00, m01, m10, tr, "R")[0]
    P = sy.cfg.P
    bulk, _ = oracle.tau2d(m00, m01, m10, ti, P, x, y)
    lay, _ = oracle.tau2d_layered(np.stack([m00] * P), np.stack([m01] * P), np.stack([m10] * P), ti, x, y)
    assert np.array_equal(bulk, lay)
