"""Energy loop over a contour with per-energy k-meshes.

Mirrors the reference's energy loop (skkr_kpar_v6/nrsp-scf/skkr.f:289-354:
zmesh contour, kmesh2d per energy with the k-set chosen from a table keyed by
Im E, t^-1 per energy, then the k-loop of cpacoord) on one device context.
The k-count grows as the energy approaches the real axis, as the reference's
k-set table does (sample/02-layer-scf: 253 k for Im E >= 0.027 up to 1953 k
at Im E = 1.1e-4).
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import observables as O
from .systems import uniform_kmesh


def geometric_mesh_sides(imag_e, nmin=256, nmax=16384):
    """Uniform-mesh side per energy: k-count geometric in Im E from nmin at the
    largest Im E to nmax at the smallest."""
    imag_e = np.asarray(imag_e, float)
    lo, hi = imag_e.min(), imag_e.max()
    sides = []
    for y in imag_e:
        t = 0.0 if hi == lo else math.log(hi / y) / math.log(hi / lo)
        nk = nmin * (nmax / nmin) ** t
        sides.append(int(round(math.sqrt(nk))))
    return sides


def run_contour(ctx, system, energies, weights, mesh_sides, reference=None,
                on_energy=None):
    """Per energy: set t^-1(E), accumulate tau_BZ over that energy's mesh.

    Returns per-energy records with tau_bz, tr G, DOS, iterations, timing, and
    the contour integral sum_E w_E tr G(E).  ``reference`` (optional) is a
    list of reference tau_bz per energy for error reports."""
    out = []
    for ie, (E, side) in enumerate(zip(energies, mesh_sides)):
        k, w = uniform_kmesh(side)
        ctx.set_tinv(*system.tinv(E))
        t0 = time.perf_counter()
        r = ctx.tau_bz(E, k, w)
        dt = time.perf_counter() - t0
        rec = {"ie": ie, "E": complex(E), "nk": len(w), "tau_bz": r["tau_bz"],
               "trG": O.tr_g(r["tau_bz"]), "dos": O.dos(r["tau_bz"]),
               "iters": r["iters"], "seconds": dt, "stats": r["stats"]}
        if reference is not None:
            rec.update(O.error_report(r["tau_bz"], reference[ie]))
        out.append(rec)
        if on_energy:
            on_energy(rec)
    total = O.integrated_g([r["trG"] for r in out], weights)
    return out, total
