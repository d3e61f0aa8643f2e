"""Write the golden vectors of phystau (screened -> physical tau).

Runs the reference's own phystau, compiled from /root/reference by
oracle/ref/Makefile into oracle/_ref/refphys_{nrsp,rsp,dbogo}:
  nrsp  : skkr_kpar_v6/nrsp-scf/phystau.f   (lm basis, n = (lmax+1)^2)
  rsp   : skkr_kpar_v6/rsp-scf/phystau.f    (kappa-mu basis, n = 2 (lmax+1)^2)
  dbogo : skkr_dbogo_bsf/rsp-scf/phystau_d.f (electron and hole halves,
          n = 4 (lmax+1)^2, screening alpha and alphah)
on seeded random inputs (lmax = 2), and stores inputs and outputs in
tests/golden/phystau.npz.

Usage: python tests/golden/make_phystau_golden.py   (needs /root/reference + flang)
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFDIR = os.path.join(ROOT, "oracle", "_ref")
LMAX = 2
DIMS = {"nrsp": (LMAX + 1) ** 2, "rsp": 2 * (LMAX + 1) ** 2, "dbogo": 4 * (LMAX + 1) ** 2}


def cases(variant, rng, ncase=2):
    n = DIMS[variant]
    out = []
    for c in range(ncase):
        tau = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        t1 = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        t2 = t1.copy() if c % 2 == 0 else rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        al = rng.normal(size=LMAX + 1) + 1j * rng.normal(size=LMAX + 1)
        ah = rng.normal(size=LMAX + 1) + 1j * rng.normal(size=LMAX + 1)
        out.append((tau, t1, t2, al, ah))
    return out


def run(variant, cs, idiag):
    n = DIMS[variant]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([len(cs), LMAX, idiag, n], np.int32).tofile(f)
            for tau, t1, t2, al, ah in cs:
                for m in (tau, t1, t2):
                    np.asfortranarray(m).T.astype(np.complex128).tofile(f)  # column-major
                al.astype(np.complex128).tofile(f)
                if variant == "dbogo":
                    ah.astype(np.complex128).tofile(f)
        subprocess.run([os.path.join(REFDIR, f"refphys_{variant}"), fin, fout], check=True)
        raw = np.fromfile(fout, np.complex128).reshape(len(cs), n, n)
        return np.ascontiguousarray(np.swapaxes(raw, 1, 2))     # back to [row, col]


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref")], check=True)
    rng = np.random.default_rng(20261203)
    data = {}
    for v in DIMS:
        for idiag in (1, 0):
            cs = cases(v, rng)
            out = run(v, cs, idiag)
            key = f"{v}_idiag{idiag}"
            data[key + "_tau"] = np.stack([c[0] for c in cs])
            data[key + "_tinv1"] = np.stack([c[1] for c in cs])
            data[key + "_tinv2"] = np.stack([c[2] for c in cs])
            data[key + "_alpha"] = np.stack([c[3] for c in cs])
            data[key + "_alphah"] = np.stack([c[4] for c in cs])
            data[key + "_out"] = out
            print(key, out.shape, np.abs(out).max())
    np.savez_compressed(os.path.join(HERE, "phystau.npz"), lmax=LMAX, **data)


if __name__ == "__main__":
    main()
