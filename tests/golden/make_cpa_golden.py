"""Write the golden vectors of the CPA solver.

Runs the reference's own tata / cpacor / cpasol, compiled from /root/reference
by oracle/ref/Makefile into oracle/_ref/refcpa_{nrsp,rsp}
(skkr_kpar_v6/nrsp-scf/fullcpa.f, rsp-scf/fullcpa.f with main/gjinv.f and
matops.f), on seeded random inputs, and stores them in tests/golden/cpa.npz:
  nrsp: n = 9 (lmax 2), 4 sites with concentrations 0.3, 1.0, 0.75 and a site
        whose mb equals ma (cpacor skips both of the last kinds);
  rsp:  n = 18, 2 sites.
Per case: the tata initial coherent t^-1, then one cpacor step from it with
the given tau_c: the new coherent t^-1, tau_A, tau_B and the error.

Usage: python tests/golden/make_cpa_golden.py   (needs /root/reference + flang)
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFDIR = os.path.join(ROOT, "oracle", "_ref")


def case(rng, n, conc):
    ns = len(conc)
    g = lambda: rng.normal(size=(ns, n, n)) + 1j * rng.normal(size=(ns, n, n))
    ma = g() + 3.0 * np.eye(n)
    mb = g() - 2.0j * np.eye(n)
    tauc = 0.3 * g()
    return np.array(conc, np.float64), ma, mb, tauc


def run(variant, cases, n):
    ns = len(cases[0][0])
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([len(cases), n, ns], np.int32).tofile(f)
            for conc, ma, mb, tauc in cases:
                conc.tofile(f)
                for arr in (ma, mb, tauc):
                    for m in arr:
                        np.asfortranarray(m).T.astype(np.complex128).tofile(f)
        subprocess.run([os.path.join(REFDIR, f"refcpa_{variant}"), fin, fout], check=True)
        raw = open(fout, "rb").read()
    out, off = [], 0
    blk = ns * n * n * 16
    for _ in cases:
        arrs = []
        for _ in range(4):
            a = np.frombuffer(raw, np.complex128, ns * n * n, off).reshape(ns, n, n)
            arrs.append(np.ascontiguousarray(np.swapaxes(a, 1, 2)))
            off += blk
        err = np.frombuffer(raw, np.float64, 1, off)[0]
        off += 8
        out.append((*arrs, err))
    assert off == len(raw)
    return out


def main():
    rng = np.random.default_rng(20261204)
    res = {}
    for variant, n, conc in (("nrsp", 9, [0.3, 1.0, 0.75, 0.5]), ("rsp", 18, [0.4, 0.9])):
        cases = [case(rng, n, conc) for _ in range(2)]
        if variant == "nrsp":
            for c in cases:
                c[2][3] = c[1][3]  # site 4: mb = ma
        out = run(variant, cases, n)
        for i, (cin, cout) in enumerate(zip(cases, out)):
            p = f"{variant}{i}_"
            res[p + "conc"], res[p + "ma"], res[p + "mb"], res[p + "tauc"] = cin
            res[p + "mc0"], res[p + "mc"], res[p + "taua"], res[p + "taub"], res[p + "error"] = cout
    path = os.path.join(HERE, "cpa.npz")
    np.savez_compressed(path, **res)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
