"""GPU off-diagonal BZ sums of site pairs (skkr_tau_bz_pairs through the
C-ABI) against the oracle (oracle.tau_pairs: the dense inverse of the finite
stack per k, pinned to the reference's tau2dgod by tests/golden/fulltau.npz).
Pairs cover a site with itself at R != 0, two sites of one principal layer,
and sites two and one PLs apart in both directions."""
import numpy as np
import pytest

import oracle
from skkr_budapest_amd import systems as S

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def gpu():
    from skkr_budapest_amd import gpu as G
    return G


@pytest.mark.parametrize("lk,s,P", [(6, 2, 3), (16, 2, 4)])
def test_pair_sums_match_oracle(gpu, lk, s, P):
    sy = S.make_system(S.small_config(lk=lk, s=s, P=P, r2max=8, tmodel="herm", energies=(0.35 + 0.04j,)))
    E = sy.cfg.energies[0]
    k, w = S.uniform_kmesh(5)
    a1, a2 = (1.0, 0.0), (0.0, 1.0)
    ns = P * s
    pairs = [(1, 1, 1, 0), (0, 1, 0, 0), (0, 2 * s + 1, 1, -1), (ns - 1, 0, 0, 2), (s, s + 1, -1, 1)]
    ref_ij, ref_ji = oracle.tau_pairs(sy, E, k, w, pairs, a1, a2)
    print(f"pair sums lk={lk} s={s} P={P} nk={len(w)}", flush=True)
    with gpu.Context.for_system(sy) as ctx:
        ctx.set_tinv(*sy.tinv(E))
        ctx.set_pairs(pairs, a1, a2)
        r = ctx.tau_bz_pairs(E, k, w)
        plain = ctx.tau_bz(E, k, w)
    for q in range(len(pairs)):
        assert rel(r["tau_ij"][q], ref_ij[q]) < 1e-11, q
        assert rel(r["tau_ji"][q], ref_ji[q]) < 1e-11, q
    np.testing.assert_array_equal(r["tau_bz"], plain["tau_bz"])
