"""The CPA oracle (oracle/cpa_oracle.c) against the reference's own
tata/cpacor/cpasol (golden vectors of tests/golden/cpa.npz, written by
tests/golden/make_cpa_golden.py from skkr_kpar_v6/{nrsp,rsp}-scf/fullcpa.f)
and against what the CPA fixes: the average-t-matrix closed form of tata, the
skipped sites, and the CPA condition c tau_A + (1-c) tau_B = tau_c at the
fixed point of a one-site lattice."""
import os

import numpy as np
import pytest

import oracle

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "cpa.npz"), allow_pickle=False)
CASES = ["nrsp0", "nrsp1", "rsp0", "rsp1"]


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("case", CASES)
def test_tata_and_cpacor_match_reference(case):
    p = case + "_"
    conc, ma, mb, tauc = GOLD[p + "conc"], GOLD[p + "ma"], GOLD[p + "mb"], GOLD[p + "tauc"]
    mc0 = np.array([oracle.tata(c, a, b) for c, a, b in zip(conc, ma, mb)])
    assert rel(mc0, GOLD[p + "mc0"]) < 1e-13
    mc, ta, tb, err = oracle.cpacor(conc, GOLD[p + "mc0"], ma, mb, tauc)
    assert rel(mc, GOLD[p + "mc"]) < 1e-11     # four chained inversions of random matrices
    assert rel(ta, GOLD[p + "taua"]) < 1e-12
    assert rel(tb, GOLD[p + "taub"]) < 1e-12
    assert abs(err - float(GOLD[p + "error"])) < 1e-12 * abs(err)


def test_tata_is_the_average_t_matrix():
    rng = np.random.default_rng(3)
    n = 6
    ma = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 3 * np.eye(n)
    mb = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) - 2j * np.eye(n)
    c = 0.35
    ref = np.linalg.inv(c * np.linalg.inv(ma) + (1 - c) * np.linalg.inv(mb))
    assert rel(oracle.tata(c, ma, mb), ref) < 1e-13


def test_skipped_sites_keep_tau_c():
    """Concentration 1 (site 2) and mb = ma (site 4 of the nrsp cases)."""
    p = "nrsp0_"
    conc, ma, mb, tauc = GOLD[p + "conc"], GOLD[p + "ma"], GOLD[p + "mb"], GOLD[p + "tauc"]
    mc, ta, tb, _ = oracle.cpacor(conc, GOLD[p + "mc0"], ma, mb, tauc)
    for li in (1, 3):
        np.testing.assert_array_equal(ta[li], tauc[li])
        np.testing.assert_array_equal(tb[li], tauc[li])
        np.testing.assert_array_equal(mc[li], GOLD[p + "mc0"][li])


def test_cpa_condition_at_the_fixed_point():
    """One-site lattice tau_c = (mc - G0)^-1: iterating cpacor converges to a
    coherent t^-1 whose projected taus average to tau_c."""
    rng = np.random.default_rng(4)
    n = 9
    G0 = 0.5 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    G0 = 0.5 * (G0 + G0.T) - 0.8j * np.eye(n)   # absorptive
    ma = np.diag(rng.uniform(-1, 1, n)) + 0.05j * np.eye(n)
    mb = np.diag(rng.uniform(-1, 1, n)) + 0.05j * np.eye(n)
    c = np.array([0.3])
    mc = oracle.tata(c[0], ma, mb)[None]
    for it in range(200):
        tauc = np.linalg.inv(mc[0] - G0)[None]
        mc, ta, tb, err = oracle.cpacor(c, mc, ma[None], mb[None], tauc)
        if err < 1e-13:
            break
    assert err < 1e-13
    tauc = np.linalg.inv(mc[0] - G0)
    assert rel(c[0] * ta[0] + (1 - c[0]) * tb[0], tauc) < 1e-10
