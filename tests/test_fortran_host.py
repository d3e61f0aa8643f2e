"""The Fortran host program (ISO_C_BINDING module skkr_gpu + the
cpacoord-style k-loop) against the oracle."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from skkr_budapest_amd import systems as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "skkr_budapest_amd", "fortran", "skkr_kloop")


def write_input(path, sy, E, k, w):
    c = sy.cfg
    with open(path, "wb") as f:
        np.array([c.lk, c.s, c.P, len(sy.R), len(w), c.itermax], np.int32).tofile(f)
        np.array([c.errmax, E.real, E.imag], np.float64).tofile(f)
        np.ascontiguousarray(sy.R, np.float64).tofile(f)
        for a in (sy.G_intra, sy.G_inter, *sy.tinv(E)):
            S.colmajor(a).tofile(f)
        np.ascontiguousarray(k, np.float64).tofile(f)
        np.ascontiguousarray(w, np.float64).tofile(f)


@pytest.mark.skipif(shutil.which("flang") is None, reason="no flang")
def test_fortran_host_builds_and_links():
    from skkr_budapest_amd import build
    exe = build.build_fortran()
    assert exe and os.path.exists(exe)


@pytest.mark.gpu
@pytest.mark.parametrize("nprocs", [1, 3])
def test_fortran_kloop_matches_oracle(tmp_path, nprocs):
    cfg = S.small_config(lk=6, s=2, P=4, r2max=4, energies=(0.4 + 0.07j,), mesh=5)
    sy = S.make_system(cfg)
    E = cfg.energies[0]
    k, w = S.uniform_kmesh(cfg.mesh)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, sy, E, k, w)
    subprocess.run([EXE, str(fin), str(fout), str(nprocs)], check=True, timeout=120)
    raw = np.fromfile(fout, np.complex128)
    ns, lk, nk = cfg.P * cfg.s, cfg.lk, len(w)
    tau = S.from_colmajor(raw[:ns * lk * lk].reshape(ns, lk, lk))
    tkd = raw[ns * lk * lk:].reshape(ns, nk).T
    ref = oracle.tau_bz(sy, E, k, w)
    assert np.linalg.norm(tau - ref["tau_bz"]) < 1e-11 * np.linalg.norm(ref["tau_bz"])
    assert np.abs(tkd - ref["trk"]).max() < 1e-11 * np.abs(ref["trk"]).max()


GEXE = os.path.join(ROOT, "skkr_budapest_amd", "fortran", "skkr_gstore")


@pytest.mark.gpu
def test_fortran_gstore_matches_reference_golden(tmp_path):
    """The producer from Fortran (skkr_gstore_main.f90 through the skkr_gpu
    module) against g2d of the reference's own gstore."""
    g = np.load(os.path.join(ROOT, "tests", "golden", "gstore.npz"), allow_pickle=False)
    p = "nb110_"
    s, lmax = int(g[p + "s"]), int(g["lmax"])
    E = complex(g[p + "E"])
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([lmax, s, int(g[p + "nextra"]), len(g[p + "ks"])], np.int32).tofile(f)
        np.array([float(g[p + "eta"]), *g[p + "a1"], *g[p + "a2"], E.real, E.imag]).tofile(f)
        np.ascontiguousarray(g[p + "cvec"], np.float64).tofile(f)
        np.ascontiguousarray(g[p + "alpha_l"], np.complex128).tofile(f)
        np.ascontiguousarray(g[p + "ks"], np.float64).tofile(f)
    subprocess.run([GEXE, str(fin), str(fout)], check=True, timeout=120)
    nl2 = (lmax + 1) ** 2
    out = np.fromfile(fout, np.complex128).reshape(len(g[p + "ks"]), 3, s, 2 * s + 1, nl2, nl2)
    for ik in range(len(g[p + "ks"])):
        ref = g[p + "g2d_L"][ik]
        assert np.abs(out[ik] - ref).max() < 1e-10 * np.abs(ref).max()
