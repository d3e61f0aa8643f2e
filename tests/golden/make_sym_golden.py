"""Write the golden vectors of the point-group symmetrisation of the BZ sum.

sphbas.npz : the matrices of every 2D point group of the reference's
    sphbas (skkr_kpar_v6/nrsp-scf/sphbas.f with matrep.f; lm basis, lmax 2):
    ng, invg, dg, dg1 per group (oracle/_ref/refsph).
sphbas_rel.npz : the same for the kappa-mu basis (rsp-scf/sphbas.f with
    matrep.f; double group, 18 x 18 for lmax 2; oracle/_ref/refsph_rel).
sym_cfg0.npz : the BZ sum of config 0's first spin channel (lk 9, 2 sites,
    8 PLs, square lattice, E = 0.5+0.05i, 16x16 mesh) symmetrised with the
    C4v matrices as nrsp-scf/cpacoord.f:281-299 does per k:
    tau_BZ += (w_k / ng) dg tau_k dg1 for every group element
    (oracle/_ref/refdrv_sym: the reference's dugo, tau2d, tripmt, addmatc).

Usage: python tests/golden/make_sym_golden.py   (needs /root/reference + flang)
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from skkr_budapest_amd import systems as S  # noqa: E402
import make_golden as G  # noqa: E402

REFDIR = os.path.join(ROOT, "oracle", "_ref")
GROUPS = ("C1", "C2", "C3", "C4", "C6", "Cs", "C2v", "C3vA", "C3vB", "C4v", "C6v")


def run_sphbas(exe="refsph"):
    with tempfile.TemporaryDirectory() as td:
        fout = os.path.join(td, "sph.bin")
        subprocess.run([os.path.join(REFDIR, exe), fout], capture_output=True, check=True)
        raw = open(fout, "rb").read()
    o = 0

    def take(dtype, n):
        nonlocal o
        a = np.frombuffer(raw[o:o + np.dtype(dtype).itemsize * n], dtype)
        o += np.dtype(dtype).itemsize * n
        return a

    lm, me, ngr = take(np.int32, 3)
    out = {"lmmax": np.int32(lm)}
    for name in GROUPS:
        ng = int(take(np.int32, 1)[0])
        invg = take(np.int32, me)[:ng]
        dg = S.from_colmajor(take(np.complex128, lm * lm * me).reshape(me, lm, lm))[:ng]
        dg1 = S.from_colmajor(take(np.complex128, lm * lm * me).reshape(me, lm, lm))[:ng]
        out[name + "_ng"] = np.int32(ng)
        out[name + "_invg"] = invg.copy()
        out[name + "_dg"] = dg
        out[name + "_dg1"] = dg1
    return out


def run_sym(sy, E, mesh, dg, dg1):
    cfg = sy.cfg
    k, w = S.uniform_kmesh(mesh)
    nk = len(w)
    tint, tl, tr = sy.tinv(E)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([cfg.lk, cfg.s, cfg.P, nk, cfg.itermax, 0], np.int32).tofile(f)
            np.array([cfg.errmax], np.float64).tofile(f)
            for t in (tint, tl, tr):
                S.colmajor(t).tofile(f)
            w.astype(np.float64).tofile(f)
            np.array([len(dg)], np.int32).tofile(f)
            S.colmajor(dg).tofile(f)
            S.colmajor(dg1).tofile(f)
            for ik in range(nk):
                gi, go = G.gk_numpy(sy, k[ik])
                S.colmajor(gi).tofile(f)
                S.colmajor(go).tofile(f)
        exe = os.path.join(REFDIR, "refdrv_sym")
        subprocess.run(["bash", "-c", f'ulimit -s unlimited; exec "{exe}" "{fin}" "{fout}"'],
                       capture_output=True, text=True, check=True)
        raw = np.fromfile(fout, np.complex128)
    ns, lk = cfg.P * cfg.s, cfg.lk
    taubz = S.from_colmajor(raw[:ns * lk * lk].reshape(ns, lk, lk))
    return dict(tau_bz=taubz, k=k, w=w, E=np.array(E), fingerprint=np.array(sy.fingerprint()))


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref")], check=True)
    sph = run_sphbas()
    np.savez_compressed(os.path.join(HERE, "sphbas.npz"), **sph)
    np.savez_compressed(os.path.join(HERE, "sphbas_rel.npz"), **run_sphbas("refsph_rel"))
    sy = S.spin_channel(S.make_system(S.CONFIGS[0]), 0)
    out = run_sym(sy, S.CONFIGS[0].energies[0], 16, sph["C4v_dg"], sph["C4v_dg1"])
    np.savez_compressed(os.path.join(HERE, "sym_cfg0.npz"), group=np.array("C4v"), **out)
    print("sphbas groups", len(GROUPS), "sym_cfg0 |tau_bz|", np.abs(out["tau_bz"]).max())


if __name__ == "__main__":
    main()
