"""The structure-constant oracle (oracle/kkrstr_oracle.c) against the
reference's own producer (golden vectors of tests/golden/gstore.npz, written
by tests/golden/make_gstore_golden.py from skkr_kpar_v6/main/gstore.f and its
callees) and against what the mathematics fixes: w(z) of scipy, the closed
form of Gamma(1/2, x), and the agreement of the two independent sums (Kambe's
Ewald split and the direct reciprocal sum) for well-separated layers."""
import os

import numpy as np
import pytest
from scipy.special import erfc, wofz

import oracle

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "gstore.npz"), allow_pickle=False)
LMAX = int(GOLD["lmax"])


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_cerf_matches_reference_on_all_branches():
    w = oracle.cerf(GOLD["cerf_z"])
    assert np.max(np.abs(w - GOLD["cerf_w"]) / np.abs(GOLD["cerf_w"])) < 1e-14


def test_cerf_is_the_faddeeva_function_to_the_references_accuracy():
    """cerf.f computes w(z) to ~1e-8 (single-precision constants in its
    Abramowitz-Stegun branch); the oracle inherits exactly that."""
    z = GOLD["cerf_z"]
    assert np.max(np.abs(oracle.cerf(z) - wofz(z)) / np.abs(wofz(z))) < 5e-8


def test_kambe_special_functions_match_reference():
    for x, ref in zip(GOLD["incgamma_in"], GOLD["incgamma_out"]):
        assert rel(oracle.incgamma(len(ref) - 1, x), ref) < 1e-13
    for (x, zz), ref in zip(GOLD["delkam_in"], GOLD["delkam_out"]):
        assert rel(oracle.delkam(len(ref) - 1, x, zz), ref) < 1e-11
    for (al, t, k), ref in zip(GOLD["ilkam_in"], GOLD["ilkam_out"]):
        assert rel(oracle.ilkam(len(ref) - 1, al, t.real, k), ref) < 1e-13


def test_incgamma_closed_form():
    """Gamma(1/2, x) = sqrt(pi) erfc(sqrt(x)) and the recursion
    Gamma(a, x) = (Gamma(a+1, x) - x^a e^-x) / a downwards, real x > 0."""
    for x in (0.3, 1.7, 4.0):
        g = oracle.incgamma(3, x)
        assert abs(g[0] - np.sqrt(np.pi) * erfc(np.sqrt(x))) < 5e-8 * abs(g[0])
        for n in range(1, 4):
            a = 0.5 - n
            assert abs(g[n] - (g[n - 1] - x ** a * np.exp(-x)) / a) < 1e-12 * abs(g[n])


@pytest.mark.parametrize("geo", ["nb110", "sq2"])
def test_lattice_kamstr_directg_match_reference(geo):
    p = geo + "_"
    L = oracle.Lattice(GOLD[p + "a1"], GOLD[p + "a2"])
    assert (L.maxr, L.maxk, L.maxshk) == tuple(int(x) for x in GOLD[p + "lat_n"][[0, 1, 3]])
    assert np.allclose([L.distr, L.distk, L.vol], GOLD[p + "lat_d"], rtol=1e-15)
    E, eta, k = complex(GOLD[p + "E"]), float(GOLD[p + "eta"]), GOLD[p + "ks"][0]
    for i, sep in enumerate(GOLD[p + "seps"]):
        alm, nsh, st = L.kamstr(LMAX, E, k, sep, eta)
        assert st == 0
        assert rel(alm, GOLD[p + "kamstr"][i]) < 1e-13
        if sep[2] > 0:
            gp, gm, ns, st = L.directg(LMAX, E, k, sep)
            assert st == 0
            assert rel(gp, GOLD[p + "directg_p"][i]) < 1e-13
            assert rel(gm, GOLD[p + "directg_m"][i]) < 1e-13


@pytest.mark.parametrize("geo", ["nb110", "sq2"])
def test_ewald_and_direct_sums_agree_for_separated_layers(geo):
    """Two independent evaluations of the same Bloch sum: Kambe's Ewald split
    (kamstr) and the plain reciprocal sum (directg, + part) must agree once
    the layers are apart; closer layers converge the direct sum slowly.  The
    bound is the reference's own accuracy: kamstr inherits ~1e-8 from cerf.f
    and the single-precision Legendre coefficients of ylmc.f:48-49, while
    directg (sphrm.f) is accurate to round-off."""
    p = geo + "_"
    L = oracle.Lattice(GOLD[p + "a1"], GOLD[p + "a2"])
    E, eta = complex(GOLD[p + "E"]), float(GOLD[p + "eta"])
    for k in GOLD[p + "ks"]:
        for sep in GOLD[p + "seps"]:
            if sep[2] < 4.5:
                continue
            alm, _, _ = L.kamstr(LMAX, E, k, sep, eta)
            gp, _, _, _ = L.directg(LMAX, E, k, sep)
            assert rel(gp, alm) < 3e-7


@pytest.mark.parametrize("geo", ["nb110", "sq2"])
def test_screened_g2d_matches_reference(geo):
    p = geo + "_"
    L = oracle.Lattice(GOLD[p + "a1"], GOLD[p + "a2"])
    s, nextra = int(GOLD[p + "s"]), int(GOLD[p + "nextra"])
    E, eta = complex(GOLD[p + "E"]), float(GOLD[p + "eta"])
    for ik, k in enumerate(GOLD[p + "ks"]):
        for side in ("L", "R"):
            al = GOLD[p + ("alpha_l" if side == "L" else "alpha_r")]
            g = L.gstore(LMAX, s, nextra, GOLD[p + "cvec"], al, E, k, eta)
            assert rel(g, GOLD[p + f"g2d_{side}"][ik]) < 1e-11


RELG = np.load(os.path.join(os.path.dirname(__file__), "golden", "relmtrx.npz"), allow_pickle=False)


def test_relmtrx_matches_reference():
    for a, fac, b in zip(RELG["a"], RELG["fac"], RELG["b"]):
        assert np.abs(oracle.relmtrx(a, fac, int(RELG["lmax"])) - b).max() < 1e-14 * np.abs(b).max()


def test_relmtrx_is_a_unitary_change_of_basis():
    """The Clebsch-Gordan matrices are orthogonal: relmtrx(I) = f I, and the
    spectrum of a becomes that of a (x) I_2 scaled by f."""
    for lmax in (1, 2, 3):
        n = (lmax + 1) ** 2
        assert np.abs(oracle.relmtrx(np.eye(n), 2.5, lmax) - 2.5 * np.eye(2 * n)).max() < 1e-14
    rng = np.random.default_rng(8)
    a = rng.normal(size=(9, 9)) + 1j * rng.normal(size=(9, 9))
    ev = np.sort_complex(np.linalg.eigvals(oracle.relmtrx(a, 1.0, 2)))
    ref = np.sort_complex(np.concatenate([np.linalg.eigvals(a)] * 2))
    assert np.abs(ev - ref).max() < 1e-12 * np.abs(ref).max()
