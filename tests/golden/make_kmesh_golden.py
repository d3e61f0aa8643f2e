"""Write the golden vectors of the 2D k-meshes (kmesh2d).

Runs the reference's own primkr2d, rgroup and kmesh2d (with speck2, ibz,
fullbz, findangles, vecrot of skkr_kpar_v6/main), compiled from
/root/reference by oracle/ref/Makefile into oracle/_ref/refkmesh, on the
cases below: every 2D lattice, irreducible wedge (intbz 1) and full BZ
(intbz 2), rotated group settings, single k-points (iset 0), the k-sets
21 and 31 of the sample 02-layer-scf (centred rectangular Nb(110), C2vx),
Cunningham special points (kunning 1) and the irreducible BZ of every group
below the lattice's holohedry.

Usage: python tests/golden/make_kmesh_golden.py   (needs /root/reference + flang)
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

A_NB = 4.667470441
# (name, ilat, idgroup, a2d, beta, delta, intbz, iset, park)
CASES = [
    ("sample_k21", 2, "C2vx", A_NB, 0.70710678120, 0.0, 1, 21, (0.0, 0.0)),
    ("sample_k31", 2, "C2vx", A_NB, 0.70710678120, 0.0, 1, 31, (0.0, 0.0)),
    ("crect_full", 2, "C2v", A_NB, 0.70710678120, 0.0, 2, 9, (0.0, 0.0)),
    ("oblique_ibz", 1, "C2", 1.0, 0.8, 0.3, 1, 8, (0.0, 0.0)),
    ("oblique_full", 1, "C2", 1.0, 0.8, 0.3, 2, 7, (0.0, 0.0)),
    ("oblique_negdelta", 1, "C2", 1.0, 0.8, -0.2, 1, 6, (0.0, 0.0)),
    ("prect_ibz", 3, "C2v", 1.0, 0.6, 0.0, 1, 10, (0.0, 0.0)),
    ("prect_full", 3, "C2vx", 1.0, 0.6, 0.0, 2, 7, (0.0, 0.0)),
    ("square_ibz", 4, "C4v", 1.0, 0.0, 0.0, 1, 12, (0.0, 0.0)),
    ("square_full", 4, "C4v", 1.0, 0.0, 0.0, 2, 9, (0.0, 0.0)),
    ("square_c2vd_full", 4, "C2vd", 1.0, 0.0, 0.0, 2, 6, (0.0, 0.0)),
    ("square_single", 4, "C4v", 1.0, 0.0, 0.0, 2, 0, (0.3, 0.1)),
    ("square_single_ibz", 4, "C4v", 1.0, 0.0, 0.0, 1, 0, (0.3, 0.1)),
    ("hex_ibz", 5, "C6v", 1.0, 0.0, 0.0, 1, 9, (0.0, 0.0)),
    ("hex_full", 5, "C6v", 1.0, 0.0, 0.0, 2, 6, (0.0, 0.0)),
    ("hex_c3va_full", 5, "C3vA", 1.0, 0.0, 0.0, 2, 5, (0.0, 0.0)),
    ("hex_single_c3vb", 5, "C3vB", 1.0, 0.0, 0.0, 2, 0, (0.4, 0.2)),
    ("hex_c2v3_full", 5, "C2v3", 1.0, 0.0, 0.0, 2, 4, (0.0, 0.0)),
]
# Cunningham special points (kunning 1: setspk, speck), with (name, ..., park, kunning)
CASES += [
    ("sample_cunn3", 2, "C2vx", A_NB, 0.70710678120, 0.0, 1, 3, (0.0, 0.0), 1),
    ("crect_cunn4_full", 2, "C2v", A_NB, 0.70710678120, 0.0, 2, 4, (0.0, 0.0), 1),
    ("prect_cunn3", 3, "C2v", 1.0, 0.6, 0.0, 1, 3, (0.0, 0.0), 1),
    ("prect_cunn2_full", 3, "C2vx", 1.0, 0.6, 0.0, 2, 2, (0.0, 0.0), 1),
    ("square_cunn4", 4, "C4v", 1.0, 0.0, 0.0, 1, 4, (0.0, 0.0), 1),
    ("square_cunn3_full", 4, "C4v", 1.0, 0.0, 0.0, 2, 3, (0.0, 0.0), 1),
    ("square_cunn_clamped", 4, "C4v", 1.0, 0.0, 0.0, 1, 9, (0.0, 0.0), 1),
    ("hex_cunn4", 5, "C6v", 1.0, 0.0, 0.0, 1, 4, (0.0, 0.0), 1),
    ("hex_cunn5_full", 5, "C6v", 1.0, 0.0, 0.0, 2, 5, (0.0, 0.0), 1),
    ("hex_cunn_clamped", 5, "C6v", 1.0, 0.0, 0.0, 1, 9, (0.0, 0.0), 1),
    ("oblique_cunn_fallback", 1, "C2", 1.0, 0.8, 0.3, 1, 5, (0.0, 0.0), 1),
    ("square_csd1_cunn3", 4, "Csd1", 1.0, 0.0, 0.0, 1, 3, (0.0, 0.0), 1),
    ("hex_c3_cunn4", 5, "C3", 1.0, 0.0, 0.0, 1, 4, (0.0, 0.0), 1),
]
# irreducible BZ of groups below the lattice's holohedry (ibz.f skip tables)
CASES += [(f"ibz_{ilat}_{g}", ilat, g, a, b, d, 1, n, (0.0, 0.0)) for ilat, a, b, d, n, groups in (
    (1, 1.0, 0.8, 0.3, 6, ("C1",)),
    (2, A_NB, 0.70710678120, 0.0, 7, ("C1", "C2", "Csx", "Csy")),
    (3, 1.0, 0.6, 0.0, 7, ("C1", "C2", "Csx", "Csy")),
    (4, 1.0, 0.0, 0.0, 8, ("C1", "C2", "C4", "C2vx", "C2vd", "Csx", "Csy", "Csd1", "Csd2")),
    (5, 1.0, 0.0, 0.0, 7, ("C1", "C2", "C3", "C6", "C3vA", "C3vB", "C2vx", "C2v2", "C2v3",
                           "Csx", "Csx2", "Csx3", "Csy", "Csy2", "Csy3")))
    for g in groups]


def run():
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([len(CASES)], np.int32).tofile(f)
            for _, ilat, grp, a, b, d, intbz, iset, park, *kun in CASES:
                np.array([ilat, intbz, kun[0] if kun else 0, iset], np.int32).tofile(f)
                np.array([a, b, d, *park], np.float64).tofile(f)
                f.write(grp.ljust(4).encode())
        subprocess.run([os.path.join(REFDIR, "refkmesh"), fin, fout], capture_output=True, check=True)
        raw = open(fout, "rb").read()
    out, o = {}, 0
    for name, *_ in CASES:
        nk = int(np.frombuffer(raw[o:o + 4], np.int32)[0]); o += 4
        arr = np.frombuffer(raw[o:o + 24 * nk], np.float64).reshape(3, nk); o += 24 * nk
        out[name + "_k"] = arr[:2].T.copy()
        out[name + "_w"] = arr[2].copy()
    return out


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref")], check=True)
    out = run()
    np.savez_compressed(os.path.join(HERE, "kmesh.npz"), **out)
    for name, *_ in CASES:
        print(f"{name:20s} nk {len(out[name + '_w'])}")


if __name__ == "__main__":
    main()
