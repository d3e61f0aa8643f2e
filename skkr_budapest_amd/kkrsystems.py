"""Layer systems on real structure constants (the layer producer of
libskkr_gpu.so, ``gpu.Structure``).

The geometry is the Nb(110) cell of the reference's sample/02-layer-scf
(input_geo.in): centred-rectangular 2D lattice, a = 4.667470441 a.u.,
beta = 0.7071067812, 4 layers per principal layer stepped by (a/2, 0, a/2),
one extra principal layer each side for the screening, Ewald parameter 1.
Screening parameters follow alphamat.f:27-49 (square well of 3 Ry, radius
0.3481599413 a, the sample's smt); the screened inverse t-matrix of each site
is that of an attractive square well (-1 Ry, same radius) minus 1/alpha, the
t-matrix form of the screening transformation, diagonal in lm.
"""
from __future__ import annotations

import numpy as np

A_NB = 4.667470441
NB110 = dict(a1=(A_NB, 0.0), a2=(0.5 * A_NB, 0.5 * A_NB * 0.70710678120), s=4, nextra=1, eta=1.0,
             cvec=[(0.5 * A_NB * i, 0.0, 0.5 * A_NB * i) for i in range(5)], rmt=0.3481599413 * A_NB)


def sph_bessel(lmax, z):
    """j_l(z), y_l(z), l = 0..lmax, complex z, by the upward recursion of
    csbf.f:21-35."""
    z = complex(z)
    j = [np.sin(z) / z, np.sin(z) / z ** 2 - np.cos(z) / z]
    y = [-np.cos(z) / z, -np.cos(z) / z ** 2 - np.sin(z) / z]
    for l in range(2, lmax + 1):
        j.append((2 * l - 1) * j[l - 1] / z - j[l - 2])
        y.append((2 * l - 1) * y[l - 1] / z - y[l - 2])
    return np.array(j[:lmax + 1]), np.array(y[:lmax + 1])


def square_well_t(E, r, v, lmax):
    """alphamat.f:27-49 (iscreen 1): t_l = R / (1 + i k R), R = -tan(delta_l)/k
    of a square well of depth v and radius r."""
    ck = np.sqrt(complex(E))
    if ck.imag < 0:
        ck = -ck
    csk = np.sqrt(complex(E - v))
    if csk.imag < 0:
        csk = -csk
    jx, yx = sph_bessel(lmax + 1, ck * r)
    jy, _ = sph_bessel(lmax + 1, csk * r)
    out = []
    for l in range(lmax + 1):
        tand = ((ck * jx[l + 1] * jy[l] - csk * jx[l] * jy[l + 1])
                / (ck * yx[l + 1] * jy[l] - csk * yx[l] * jy[l + 1]))
        react = -tand / ck
        out.append(react / (1.0 + 1j * ck * react))
    return np.array(out)


def clebsch_kappa_l(lmax):
    """l of each kappa-mu index (2 (lmax+1)^2 of them) in the order of
    clebsch.f, ir = 2 kap^2 + kap + m (+1): kap = -1, 1, -2, 2, -3, ..."""
    order = []
    for l in range(lmax + 1):
        if l > 0:
            order += [l] * (2 * l)        # kap = l, j = l - 1/2
        order += [l] * (2 * l + 2)        # kap = -l - 1, j = l + 1/2
    return order


def nb110_workload(E, P=3, lmax=2, geo=NB110, rel=0):
    """Producer arguments, screening alpha[layer, l] and the screened t^-1 of
    the interface (P*s sites) and of the leads (s sites): (lmax+1)^2 blocks,
    or 2 (lmax+1)^2 kappa-mu blocks for rel (t^-1 depends on l only, so it
    stays diagonal)."""
    s = geo["s"]
    al = np.array([square_well_t(E, geo["rmt"], 3.0, lmax) for _ in range(s)])
    tv = 1.0 / square_well_t(E, geo["rmt"], -1.0, lmax) - 1.0 / al[0]
    lof = clebsch_kappa_l(lmax) if rel else [l for l in range(lmax + 1) for _ in range(2 * l + 1)]
    ti = np.array([np.diag([tv[l] for l in lof]) for _ in range(s)])
    return dict(structure=dict(lmax=lmax, s=s, nextra=geo["nextra"], eta=geo["eta"], a1=geo["a1"],
                               a2=geo["a2"], cvec=np.array(geo["cvec"])),
                alpha=al, tinv=np.concatenate([ti] * P), tinv_lead=ti, P=P)


def reciprocal_mesh(a1, a2, n):
    """n x n Monkhorst-Pack points of the reciprocal cell, equal weights."""
    a1, a2 = np.asarray(a1, float), np.asarray(a2, float)
    det = a1[0] * a2[1] - a1[1] * a2[0]
    b1 = 2 * np.pi * np.array([a2[1], -a2[0]]) / det
    b2 = 2 * np.pi * np.array([-a1[1], a1[0]]) / det
    f = (np.arange(n) + 0.5) / n - 0.5
    k = np.array([u * b1 + v * b2 for u in f for v in f])
    return k, np.full(len(k), 1.0 / len(k))
