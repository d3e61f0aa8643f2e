"""Write golden vectors of the lm -> kappa-mu transformation of structure
constants: the reference's clebsch + relmtrx (skkr_kpar_v6/main/clebsch.f,
relmtrx.f, compiled where they lie by oracle/ref/Makefile into
oracle/_ref/refrel) on seeded random 9 x 9 matrices (lmax 2) with complex
factors, stored in tests/golden/relmtrx.npz.

Usage: python tests/golden/make_relmtrx_golden.py   (needs /root/reference + flang)
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    rng = np.random.default_rng(20261205)
    lmax, nc = 2, 4
    a = rng.normal(size=(nc, 9, 9)) + 1j * rng.normal(size=(nc, 9, 9))
    fac = 1.0 + 1e-5 * (rng.normal(size=nc) + 1j * rng.normal(size=nc))
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([nc, lmax], np.int32).tofile(f)
            for c in range(nc):
                np.array([fac[c]], np.complex128).tofile(f)
                np.asfortranarray(a[c]).T.astype(np.complex128).tofile(f)
        subprocess.run([os.path.join(ROOT, "oracle", "_ref", "refrel"), fin, fout], check=True)
        b = np.fromfile(fout, np.complex128).reshape(nc, 18, 18)
    b = np.ascontiguousarray(np.swapaxes(b, 1, 2))
    path = os.path.join(HERE, "relmtrx.npz")
    np.savez_compressed(path, a=a, fac=fac, b=b, lmax=lmax)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
