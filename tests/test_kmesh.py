"""The 2D k-meshes (skkr_budapest_amd.kmesh, a restatement of kmesh2d.f)
against the reference's own primkr2d/rgroup/kmesh2d (tests/golden/kmesh.npz)
and the k-counts the sample prints."""
import os
import sys

import numpy as np
import pytest

from skkr_budapest_amd import kmesh as KM

GOLD = os.path.join(os.path.dirname(__file__), "golden")
sys.path.insert(0, GOLD)
from make_kmesh_golden import A_NB, CASES  # noqa: E402


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_kmesh_matches_reference(case):
    g = np.load(os.path.join(GOLD, "kmesh.npz"))
    name, ilat, grp, a, b, d, intbz, iset, park = case
    k, w = KM.kmesh2d(intbz, iset, KM.primkr2d(ilat, grp, a, b, d), park)
    kr, wr = g[name + "_k"], g[name + "_w"]
    assert k.shape == kr.shape
    assert np.abs(k - kr).max() <= 2e-15 * max(1.0, np.abs(kr).max())
    assert np.abs(w - wr).max() <= 1e-16
    assert abs(w.sum() - 1.0) < 1e-14


def test_sample_k_sets_have_the_printed_sizes():
    """sample/02-layer-scf: KSETTAB 21, 31, 41, 51, 61 (input_rsp.in:35-39)
    with intbz 1 on the C2vx centred-rectangular Nb(110) lattice print
    nk = 253, 528, 903, 1378, 1953 (Nb8_110.prn:268-298)."""
    lat = KM.primkr2d(2, "C2vx", A_NB, 0.70710678120, 0.0)
    for iset, nk in ((21, 253), (31, 528), (41, 903), (51, 1378), (61, 1953)):
        k, w = KM.kmesh2d(1, iset, lat)
        assert len(w) == nk


def test_unrestated_settings_raise():
    with pytest.raises(NotImplementedError):   # wedge of a subgroup (ibz.f tables)
        KM.kmesh2d(1, 8, KM.primkr2d(4, "C2vx", 1.0))
    with pytest.raises(NotImplementedError):   # Cunningham points
        KM.kmesh2d(1, 8, KM.primkr2d(4, "C4v", 1.0), kunning=1)
    with pytest.raises(ValueError):
        KM.primkr2d(2, "C4v", 1.0, 0.7)
