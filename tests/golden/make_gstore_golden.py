"""Write the golden vectors of the layer structure constants.

Runs the reference's own producer, compiled from /root/reference by
oracle/ref/Makefile into oracle/_ref/refgstore (skkr_kpar_v6/main/gstore.f
with kamstr.f, directg.f, scrstr.f, scrinv.f, cerf.f, incgamma.f, delkam.f,
ilkam.f, ylmc.f, gauntcalc.f, blm.f, sphrm.f and the 2D lattice routines),
and stores inputs and outputs in tests/golden/gstore.npz:
  - w(z) of cerf.f on 305 points of all its branches; incgamma.f, delkam.f,
    ilkam.f on 10 random arguments each;
  - two geometries:
      "nb110": the centred-rectangular Nb(110) cell of
               sample/02-layer-scf/input_geo.in (a = 4.667470441 a.u.,
               beta = 0.7071067812), 4 layers per PL stepped by (a/2, 0, a/2)
               as in that sample's layer list, 1 extra PL for screening,
               E = 0.5 + 0.05i;
      "sq2":   a square cell (a = 5.2 a.u.), 2 layers per PL stepped by
               (a/2, a/2, 2.6), 1 extra PL, E = -0.2 + 0.01i;
    with screening parameters of the square-well form of alphamat.f:16-49
    (V = 3 Ry, well radius 0.3481599413 a), eta = 1;
  - per geometry: the lattice sizes of vector2d.f, kamstr.f and directg.f
    for a list of layer separations at the first k, and g2d of gstore.f
    (bulk geometry, sides 'L' and 'R') at two k.

Usage: python tests/golden/make_gstore_golden.py   (needs /root/reference + flang)
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np
from scipy.special import spherical_jn, spherical_yn

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "refgstore")
LMAX = 2


def square_well_alpha(E, r, lmax=LMAX, v=3.0):
    """alphamat.f:27-49 (iscreen = 1) with scipy's spherical Bessel functions."""
    ck = np.sqrt(complex(E))
    if ck.imag < 0:
        ck = -ck
    csk = np.sqrt(complex(E - v))
    if csk.imag < 0:
        csk = -csk
    x, y = ck * r, csk * r
    out = []
    for l in range(lmax + 1):
        tand = ((ck * spherical_jn(l + 1, x) * spherical_jn(l, y) - csk * spherical_jn(l, x) * spherical_jn(l + 1, y))
                / (ck * spherical_yn(l + 1, x) * spherical_jn(l, y) - csk * spherical_yn(l, x) * spherical_jn(l + 1, y)))
        react = -tand / ck
        out.append(react / (1.0 + 1j * ck * react))
    return np.array(out)


def geometries():
    a = 4.667470441
    nb = dict(name="nb110", a1=(a, 0.0), a2=(0.5 * a, 0.5 * a * 0.70710678120), s=4, nextra=1,
              cvec=[(0.5 * a * i, 0.0, 0.5 * a * i) for i in range(5)], E=complex(0.5, 0.05),
              eta=1.0, rmt=0.3481599413 * a,
              seps=[(0, 0, 0), (0.5 * a, 0, 0.5 * a), (-0.5 * a, 0, -0.5 * a), (a, 0, a),
                    (1.5 * a, 0, 1.5 * a), (0.3, 0.2, 0.0), (a, 0.5, 0.0)],
              ks=[(0.1, 0.2), (0.3, -0.1)])
    b = 5.2
    sq = dict(name="sq2", a1=(b, 0.0), a2=(0.0, b), s=2, nextra=1,
              cvec=[(0.5 * b * i, 0.5 * b * i, 2.6 * i) for i in range(3)], E=complex(-0.2, 0.01),
              eta=1.0, rmt=0.3481599413 * b,
              seps=[(0, 0, 0), (0.5 * b, 0.5 * b, 2.6), (b, b, 5.2), (1.5 * b, 1.5 * b, 7.8)],
              ks=[(0.05, 0.4), (-0.6, 0.11)])
    return [nb, sq]


def run(fin_bytes):
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(fin_bytes)
        subprocess.run([REF, fin, fout], check=True)
        with open(fout, "rb") as f:
            return f.read()


def special(rng):
    z = np.concatenate([rng.uniform(-6, 6, 200) + 1j * rng.uniform(-6, 6, 200),
                        rng.uniform(-1.5, 1.5, 100) + 1j * rng.uniform(-0.99, 0.99, 100),
                        [0.0, 1e-12, 3.9 + 0.5j, -4.1 - 0.2j, 2 + 1j]])
    gi = [(4, complex(*rng.uniform(-3, 3, 2))) for _ in range(10)]
    di = [(4, complex(*rng.uniform(-3, 3, 2)), complex(*rng.uniform(-3, 3, 2))) for _ in range(10)]
    ii = [(4, complex(rng.uniform(0.1, 1), rng.uniform(-.2, .2)), rng.uniform(1, 10),
           complex(rng.uniform(0.3, 1), rng.uniform(0, .2))) for _ in range(10)]
    buf = [np.array([1, len(z)], np.int32).tobytes(), z.astype(np.complex128).tobytes(),
           np.array([len(gi)], np.int32).tobytes()]
    for n, x in gi:
        buf += [np.array([n], np.int32).tobytes(), np.array([x], np.complex128).tobytes()]
    buf.append(np.array([len(di)], np.int32).tobytes())
    for n, x, zz in di:
        buf += [np.array([n], np.int32).tobytes(), np.array([x, zz], np.complex128).tobytes()]
    buf.append(np.array([len(ii)], np.int32).tobytes())
    for l, al, t, k in ii:
        buf += [np.array([l], np.int32).tobytes(), np.array([al], np.complex128).tobytes(),
                np.array([t]).tobytes(), np.array([k], np.complex128).tobytes()]
    o = np.frombuffer(run(b"".join(buf)), np.complex128)
    out = {"cerf_z": z, "cerf_w": o[:len(z)]}
    off = len(z)
    for tag, cases in (("incgamma", gi), ("delkam", di), ("ilkam", ii)):
        res = []
        for c in cases:
            res.append(o[off:off + c[0] + 1])
            off += c[0] + 1
        out[f"{tag}_out"] = np.array(res)
    assert off == len(o)
    out["incgamma_in"] = np.array([x for _, x in gi])
    out["delkam_in"] = np.array([[x, zz] for _, x, zz in di])
    out["ilkam_in"] = np.array([[al, t, k] for _, al, t, k in ii])
    return out


def structure(g):
    s, nl2 = g["s"], (LMAX + 1) ** 2
    al = np.array([square_well_alpha(g["E"], g["rmt"]) for _ in range(s)])
    ar = np.array([square_well_alpha(g["E"], g["rmt"] * 0.98) for _ in range(s)])
    buf = [np.array([2, LMAX, s, g["nextra"], len(g["seps"]), len(g["ks"])], np.int32).tobytes(),
           np.array([g["eta"], *g["a1"], *g["a2"], g["E"].real, g["E"].imag]).tobytes(),
           np.array(g["cvec"], np.float64).tobytes(), al.astype(np.complex128).tobytes(),
           ar.astype(np.complex128).tobytes(), np.array(g["seps"], np.float64).tobytes(),
           np.array(g["ks"], np.float64).tobytes()]
    raw = run(b"".join(buf))
    hdr = np.frombuffer(raw, np.int32, 4, 0)
    dd = np.frombuffer(raw, np.float64, 3, 16)
    c = np.frombuffer(raw, np.complex128, offset=40)
    off, kam, dgp, dgm, gL, gR = 0, [], [], [], [], []
    for _ in g["seps"]:
        for lst in (kam, dgp, dgm):
            lst.append(c[off:off + nl2 * nl2].reshape(nl2, nl2).T)
            off += nl2 * nl2
    n = nl2 * nl2 * (2 * s + 1) * s * 3
    for _ in g["ks"]:
        for lst in (gL, gR):
            lst.append(c[off:off + n].reshape(3, s, 2 * s + 1, nl2, nl2))
            off += n
    assert off == len(c)
    p = g["name"] + "_"
    return {p + "a1": np.array(g["a1"]), p + "a2": np.array(g["a2"]), p + "s": s, p + "nextra": g["nextra"],
            p + "cvec": np.array(g["cvec"]), p + "E": g["E"], p + "eta": g["eta"], p + "alpha_l": al,
            p + "alpha_r": ar, p + "seps": np.array(g["seps"], np.float64), p + "ks": np.array(g["ks"]),
            p + "lat_n": hdr.copy(), p + "lat_d": dd.copy(), p + "kamstr": np.array(kam),
            p + "directg_p": np.array(dgp), p + "directg_m": np.array(dgm), p + "g2d_L": np.array(gL),
            p + "g2d_R": np.array(gR)}


def main():
    rng = np.random.default_rng(20261203)
    out = special(rng)
    for g in geometries():
        out.update(structure(g))
    out["lmax"] = LMAX
    out["geometries"] = np.array([g["name"] for g in geometries()])
    path = os.path.join(HERE, "gstore.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
