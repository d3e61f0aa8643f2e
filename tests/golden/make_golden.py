"""Write the golden vectors of the per-(E,k) layer kernel.

Runs the reference's own subroutines (dugo, surfgfd, packer, tau2d, gjinv,
matops), compiled from /root/reference by oracle/ref/Makefile into
oracle/_ref/, on systems produced by skkr_budapest_amd.systems, and stores
inputs' fingerprints and outputs as small .npz fixtures here.

Cases
  cfg0_spin1, cfg0_spin2 : the two nonrelativistic spin channels of config 0
      (lk=9, 2 sites/PL, 8 PLs, 13 R, E=0.5+0.05i, full 16x16 mesh), through
      the kpar subroutines -- the reference runs spin channels separately
      (skkr_kpar_v6/nrsp-scf/skkr.f:258).
  bdg18 : dense e/h-doubled 18x18 site blocks (lk=18, 2 sites/PL, 4 PLs,
      13 R, pairing 0.05), 6x6 mesh, through the dbogo twins.
  herm16 : dense Hermitian-model 16x16 site blocks (4 sites/PL, 4 PLs, 25 R),
      4x4 mesh, through the dbogo twins.

Usage: python tests/golden/make_golden.py   (needs /root/reference + flang)
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from skkr_budapest_amd import systems as S  # noqa: E402

REFDIR = os.path.join(ROOT, "oracle", "_ref")


def golden_cases():
    c0 = S.make_system(S.CONFIGS[0])
    cases = {
        "cfg0_spin1": ("kpar", S.spin_channel(c0, 0), S.CONFIGS[0].energies[0], 16),
        "cfg0_spin2": ("kpar", S.spin_channel(c0, 1), S.CONFIGS[0].energies[0], 16),
    }
    cb = S.Config("golden-bdg18", 71, lk=18, s=2, P=4, r2max=4, tmodel="bdg",
                  delta=0.05, energies=(0.35 + 0.06j,), mesh=6)
    cases["bdg18"] = ("dbogo", S.make_system(cb), cb.energies[0], cb.mesh)
    ch = S.Config("golden-herm16", 72, lk=16, s=4, P=4, r2max=8, tmodel="herm",
                  energies=(0.4 + 0.08j,), mesh=4)
    cases["herm16"] = ("dbogo", S.make_system(ch), ch.energies[0], ch.mesh)
    return cases


def gk_numpy(sy, k):
    ph = np.exp(-1j * (sy.R @ k))
    gi = np.tensordot(ph, sy.G_intra, axes=(0, 0))
    go = np.tensordot(ph, sy.G_inter, axes=(0, 0))
    return gi, go


def save_indices(nk):
    return sorted({0, nk // 3, nk - 1})


def run_case(fam, sy, E, mesh):
    cfg = sy.cfg
    k, w = S.uniform_kmesh(mesh)
    nk = len(w)
    isave = save_indices(nk)
    tint, tl, tr = sy.tinv(E)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([cfg.lk, cfg.s, cfg.P, nk, cfg.itermax, len(isave)], np.int32).tofile(f)
            np.array([cfg.errmax], np.float64).tofile(f)
            np.array([i + 1 for i in isave], np.int32).tofile(f)
            for t in (tint, tl, tr):
                S.colmajor(t).tofile(f)
            w.astype(np.float64).tofile(f)
            for ik in range(nk):
                gi, go = gk_numpy(sy, k[ik])
                S.colmajor(gi).tofile(f)
                S.colmajor(go).tofile(f)
        exe = os.path.join(REFDIR, f"refdrv_{fam}")
        proc = subprocess.run(["bash", "-c", f'ulimit -s unlimited; exec "{exe}" "{fin}" "{fout}"'],
                              capture_output=True, text=True, check=True)
        raw = np.fromfile(fout, np.complex128)
    ns, lk = cfg.P * cfg.s, cfg.lk
    o = 0
    taubz = S.from_colmajor(raw[o:o + ns * lk * lk].reshape(ns, lk, lk)); o += ns * lk * lk
    trk = raw[o:o + ns * nk].reshape(nk, ns); o += ns * nk
    tsave = S.from_colmajor(raw[o:o + len(isave) * ns * lk * lk].reshape(len(isave), ns, lk, lk))
    # surfgfd prints " iter=   N  error=..." per iteration after " Side: L|R"
    iters, side, cur = [], None, 0
    for line in proc.stdout.splitlines():
        m = re.match(r"\s*Side:\s*([LR])", line)
        if m:
            if side is not None:
                iters.append(cur)
            side, cur = m.group(1), 0
            continue
        m = re.match(r"\s*iter=\s*(\d+)", line)
        if m:
            cur = int(m.group(1))
    if side is not None:
        iters.append(cur)
    iters = np.array(iters, np.int32).reshape(nk, 2)
    return dict(tau_bz=taubz, trk=trk, tau_save=tsave, save_idx=np.array(isave),
                iters=iters, k=k, w=w, E=np.array(E), fingerprint=np.array(sy.fingerprint()),
                lk=cfg.lk, s=cfg.s, P=cfg.P, errmax=cfg.errmax, itermax=cfg.itermax)


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref")], check=True)
    for name, (fam, sy, E, mesh) in golden_cases().items():
        out = run_case(fam, sy, E, mesh)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(name, fam, "nk", len(out["w"]), "iters L/R mean",
              out["iters"].mean(0), "|tau_bz|", np.abs(out["tau_bz"]).max())


if __name__ == "__main__":
    main()
