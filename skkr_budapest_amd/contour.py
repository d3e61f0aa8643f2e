"""Energy contour of the k-loop's energy loop.

Restates the semicircle branch (imesh=1) of zmesh
(skkr_kpar_v6/main/zmesh.f:55-83) with Gauss-Legendre nodes from legzero
(skkr_kpar_v6/main/legzero.f:2-40) and legpol (skkr_kpar_v6/main/legpol.f).
With a log mesh parameter epsilon > 1e-5 the angle is mapped as
phi = eps*(exp(-y)-1), eps=1/epsilon, which crowds the nodes towards E_F.
"""
from __future__ import annotations

import math

import numpy as np


def legpol(x: float, l: int):
    """P_l(x), P_{l-1}(x) and P_l'(x) by the three-term recursion."""
    pnm1, pn = 1.0, x
    for n in range(1, l):
        pnp1 = ((n + n + 1) * x * pn - n * pnm1) / (n + 1.0)
        pnm1, pn = pn, pnp1
    dpl = l * (x * pn - pnm1) / (x * x - 1.0)
    return pn, pnm1, dpl


def legzero(ng: int):
    """Gauss-Legendre nodes (ascending) and weights on [-1, 1]; Newton from the
    Abramowitz-Stegun starting values, stop at |dx| <= 1e-14."""
    xx = np.zeros(ng)
    wx = np.zeros(ng)
    n = ng // 2
    for i in range(1, n + 1):
        phi = (4 * i - 1) * math.pi / (4 * ng + 2.0)
        x = math.cos(phi + 1.0 / (8 * ng * ng * math.tan(phi)))
        for _ in range(100):
            pl, plm1, dpl = legpol(x, ng)
            dx = pl / dpl
            x = x - dx
            if abs(dx) <= 1e-14:
                break
        else:
            raise RuntimeError("legzero: Newton did not converge")
        pl, plm1, dpl = legpol(x, ng)
        w = 2.0 * (1.0 - x * x) / (ng * ng * plm1 * plm1)
        xx[ng - i] = x
        xx[i - 1] = -x
        wx[ng - i] = w
        wx[i - 1] = w
    if 2 * n != ng:
        pl, plm1, dpl = legpol(0.0, ng)
        xx[n] = 0.0
        wx[n] = 2.0 / (ng * ng * plm1 * plm1)
    return xx, wx


def semicircle(ne: int, ebottom: float, etop: float, epsilon: float):
    """Contour energies z and weights ww (zmesh.f:55-83, one panel)."""
    eps0 = 1e-5
    r = (etop - ebottom) / 2.0
    phi1, phi2 = math.pi, 0.0
    z0 = ebottom + r
    x, w = legzero(ne)
    z = np.zeros(ne, np.complex128)
    ww = np.zeros(ne, np.complex128)
    if epsilon > eps0:
        eps = 1.0 / epsilon
        y1 = -math.log(1.0 + phi1 / eps)
        y2 = -math.log(1.0 + phi2 / eps)
        for i in range(ne):
            y = 0.5 * (y2 - y1) * x[i] + 0.5 * (y2 + y1)
            phi = eps * (math.exp(-y) - 1.0)
            z[i] = r * np.exp(1j * phi) + z0
            ww[i] = -0.5 * (y2 - y1) * eps * math.exp(-y) * 1j * (z[i] - z0) * w[i]
    else:
        for i in range(ne):
            phi = 0.5 * (phi2 - phi1) * x[i] + 0.5 * (phi2 + phi1)
            z[i] = r * np.exp(1j * phi) + z0
            ww[i] = 0.5 * (phi2 - phi1) * 1j * (z[i] - z0) * w[i]
    return z, ww


# Contour of the 16-energy configurations: E_b=-0.8, E_F=0.6, radius 0.7; the
# log-mesh parameter puts the node nearest E_F at Im E = 0.005 Ry.
CONTOUR_EB = -0.8
CONTOUR_EF = 0.6
CONTOUR_NE = 16


def epsilon_for_min_imag(ne, ebottom, etop, im_min):
    lo, hi = 1e-3, 1e4
    for _ in range(200):
        mid = math.sqrt(lo * hi)
        z, _ = semicircle(ne, ebottom, etop, mid)
        if z.imag.min() > im_min:
            lo = mid
        else:
            hi = mid
    return math.sqrt(lo * hi)


def config_contour():
    eps = epsilon_for_min_imag(CONTOUR_NE, CONTOUR_EB, CONTOUR_EF, 0.005)
    return semicircle(CONTOUR_NE, CONTOUR_EB, CONTOUR_EF, eps)
