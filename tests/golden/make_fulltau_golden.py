"""Write golden vectors of the full tau (every pair of interface layers).

Runs the reference's tau2dgod (skkr_dbogo_bsf/clu/tau2d_god.f, the Godfrin
recursion extended to all layer pairs that the cluster embedding reads),
compiled with the dbogo main/ routines by oracle/ref/Makefile into
oracle/_ref/refdrv_god, on a dense Hermitian-model system (lk = 8, 2 sites per
PL, 3 PLs, 13 R, E = 0.4 + 0.06i) for every k of a 3 x 3 mesh (weight factor
1), and stores the full tau per k in tests/golden/fulltau.npz.

Usage: python tests/golden/make_fulltau_golden.py   (needs /root/reference + flang)
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
from make_golden import gk_numpy  # noqa: E402

REFDIR = os.path.join(ROOT, "oracle", "_ref")


def main():
    cfg = S.Config("golden-fulltau", 73, lk=8, s=2, P=3, r2max=4, tmodel="herm",
                   energies=(0.4 + 0.06j,), mesh=3)
    sy = S.make_system(cfg)
    E = cfg.energies[0]
    k, w = S.uniform_kmesh(cfg.mesh)
    nk = len(w)
    tint, tl, tr = sy.tinv(E)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([cfg.lk, cfg.s, cfg.P, nk, cfg.itermax, nk], np.int32).tofile(f)
            np.array([cfg.errmax], np.float64).tofile(f)
            np.arange(1, nk + 1, dtype=np.int32).tofile(f)
            for t in (tint, tl, tr):
                S.colmajor(t).tofile(f)
            w.astype(np.float64).tofile(f)
            for ik in range(nk):
                gi, go = gk_numpy(sy, k[ik])
                S.colmajor(gi).tofile(f)
                S.colmajor(go).tofile(f)
        exe = os.path.join(REFDIR, "refdrv_god")
        subprocess.run(["bash", "-c", f'ulimit -s unlimited; exec "{exe}" "{fin}" "{fout}"'],
                       capture_output=True, text=True, check=True)
        raw = np.fromfile(fout, np.complex128)
    ns, lk = cfg.P * cfg.s, cfg.lk
    o = ns * lk * lk + ns * nk  # taubz, trk
    full = raw[o:].reshape(nk, ns, ns, lk, lk)          # [k][lj][li][b][a]
    full = np.ascontiguousarray(np.transpose(full, (0, 2, 1, 4, 3)))  # [k][li][lj][a][b]
    path = os.path.join(HERE, "fulltau.npz")
    np.savez_compressed(path, full_tau=full, k=k, w=w, E=np.array(E), fingerprint=np.array(sy.fingerprint()),
                        lk=cfg.lk, s=cfg.s, P=cfg.P, seed=73, r2max=4)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
