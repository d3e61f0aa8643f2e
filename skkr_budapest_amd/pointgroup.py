"""Matrix representations of the 2D point groups in the lm and kappa-mu bases:
the dg, dg1, ng, invg that the reference's sphbas computes
(skkr_kpar_v6/nrsp-scf/sphbas.f:2 for the lm basis, rsp-scf/sphbas.f:2 for
the kappa-mu basis, both with matrep.f) and that the BZ sum of an
irreducible wedge applies per k (nrsp-scf/cpacoord.f:294-299;
skkr_set_point_group).

The reference builds each matrix from irreducible-representation tables
(Altmann & Herzig) and a symmetry-adapted basis.  Here they come from the
operations themselves:
- rotation "R" phi about z: D = diag(e^{-i m phi}), which takes f(r) to
  f(R(-phi) r);
- mirror "M" alpha: D[(l, m), (l, -m)] = (-1)^m e^{-2 i m alpha}, which takes
  f(theta, phi) to f(theta, -2 alpha - phi), the reflection through the
  vertical plane at azimuth -alpha;
- kappa-mu basis: C (D_lm (x) U) C^T, with C the Clebsch-Gordan matrix of
  relmtrx.f / clebsch.f (|kappa mu>, kappa = -1, 1, -2, 2, ...; spin order
  ms = -1/2, +1/2), U = diag(e^{i phi/2}, e^{-i phi/2}) for a rotation by
  phi in (-pi, pi] and, for a mirror, U = -i [[0, e^{i b}], [e^{-i b}, 0]]
  in that spin order with b = alpha + pi/2.

What has to be alike is the reference's own order of the group elements
and, for the kappa-mu mirrors, which of the two normals +-n it takes (the
double group's sign): those are the table below.  Its Cs in the kappa-mu
basis is built from the single-valued characters +-1 of Cs
(rsp-scf/sphbas.f:336-499), so its mirror squares to +1: the double-group
mirror of the xz plane with the rows of mu > 0 negated.
"""
from __future__ import annotations

import math

import numpy as np

GROUPS = ("C1", "C2", "C3", "C4", "C6", "Cs", "C2v", "C3vA", "C3vB", "C4v", "C6v")

# (kind, angle in units of pi/12, kappa-mu sign): "R" rotation by the angle
# (in (-pi, pi]), "M" the mirror of that parameter; in sphbas's order.
_T = {
    "C1": [("R", 0, 1)],
    "C2": [("R", 0, 1), ("R", 12, 1)],
    "C3": [("R", 0, 1), ("R", 8, 1), ("R", -8, 1)],
    "C4": [("R", 0, 1), ("R", 6, 1), ("R", -6, 1), ("R", 12, 1)],
    "C6": [("R", 0, 1), ("R", 4, 1), ("R", -4, 1), ("R", 8, 1), ("R", -8, 1), ("R", 12, 1)],
    "Cs": [("R", 0, 1), ("Ms", 0, 1)],
    "C2v": [("R", 0, 1), ("R", 12, 1), ("M", 6, -1), ("M", 0, 1)],
    "C3vA": [("R", 0, 1), ("R", 8, 1), ("R", -8, 1), ("M", 0, 1), ("M", 8, 1), ("M", 4, -1)],
    "C3vB": [("R", 0, 1), ("R", 8, 1), ("R", -8, 1), ("M", 6, -1), ("M", 2, 1), ("M", 10, 1)],
    "C4v": [("R", 0, 1), ("R", 6, 1), ("R", -6, 1), ("R", 12, 1), ("M", 6, -1), ("M", 0, 1),
            ("M", 9, -1), ("M", 3, 1)],
    "C6v": [("R", 0, 1), ("R", 4, 1), ("R", -4, 1), ("R", 8, 1), ("R", -8, 1), ("R", 12, 1),
            ("M", 6, -1), ("M", 2, 1), ("M", 10, 1), ("M", 0, 1), ("M", 8, 1), ("M", 4, -1)],
}


def _lm(lmax):
    return [(l, m) for l in range(lmax + 1) for m in range(-l, l + 1)]


def lm_matrix(kind, angle, lmax):
    """The operation on Y_lm, l <= lmax (rows and columns in lm order)."""
    idx = _lm(lmax)
    n = len(idx)
    if kind == "R":
        return np.diag([np.exp(-1j * m * angle) for _, m in idx])
    d = np.zeros((n, n), np.complex128)
    for a, (l, m) in enumerate(idx):
        d[a, idx.index((l, -m))] = (-1) ** (m % 2) * np.exp(-2j * m * angle)
    return d


def clebsch_matrix(lmax):
    """C[kappa-mu index, 2 lm + s] (s = 0: ms = -1/2, s = 1: +1/2): the
    real orthogonal matrix with relmtrx(a) = C (a (x) 1_2) C^T."""
    idx = _lm(lmax)
    n = len(idx)
    c = np.zeros((2 * n, 2 * n))
    r = 0
    # by j, and for each j kappa = -(j+1/2) (l = j - 1/2) before kappa = j+1/2
    # (l = j + 1/2): kappa = -1, 1, -2, 2, ...; mu from -j to j
    for j2 in range(1, 2 * lmax + 2, 2):          # 2j = 1, 3, 5, ...
        for l in ((j2 - 1) // 2, (j2 - 1) // 2 + 1):
            if l > lmax:
                continue
            for mu2 in range(-j2, j2 + 1, 2):         # 2 mu
                mu = mu2 / 2.0
                t = 2 * l + 1
                if j2 == 2 * l + 1:   # j = l + 1/2
                    cd = math.sqrt((l - mu + 0.5) / t)   # |m = mu + 1/2, down>
                    cu = math.sqrt((l + mu + 0.5) / t)   # |m = mu - 1/2, up>
                else:                 # j = l - 1/2
                    cd = math.sqrt((l + mu + 0.5) / t)
                    cu = -math.sqrt((l - mu + 0.5) / t)
                md, mu_ = int(round(mu + 0.5)), int(round(mu - 0.5))
                if abs(md) <= l:
                    c[r, 2 * idx.index((l, md)) + 0] = cd
                if abs(mu_) <= l:
                    c[r, 2 * idx.index((l, mu_)) + 1] = cu
                r += 1
    return c


def _spin(kind, angle):
    if kind == "R":
        return np.diag([np.exp(0.5j * angle), np.exp(-0.5j * angle)])
    b = angle + 0.5 * math.pi
    return -1j * np.array([[0.0, np.exp(1j * b)], [np.exp(-1j * b), 0.0]])


def group_matrices(group: str, lmax: int = 2, rel: bool = False):
    """(dg, dg1, invg) of a group as sphbas returns them: dg, dg1 (ng, n, n)
    [row, col], n = (lmax+1)^2 (lm) or 2 (lmax+1)^2 (kappa-mu); invg (ng,)
    1-based index of each element's inverse."""
    if group not in _T:
        raise ValueError(f"unknown group {group!r}; one of {GROUPS}")
    els = _T[group]
    c = clebsch_matrix(lmax) if rel else None
    dg = []
    for kind, k, sgn in els:
        ang = k * math.pi / 12.0
        base = "R" if kind == "R" else "M"
        d = lm_matrix(base, ang, lmax)
        if rel:
            d = c @ np.kron(d, _spin(base, ang)) @ c.T
            if kind == "Ms":
                # single-valued characters of Cs: negate the rows of mu > 0
                mus = _mus(lmax)
                d = np.where((mus > 0)[:, None], -d, d)
            else:
                d = sgn * d
        dg.append(d)
    dg = np.array(dg)
    dg1 = np.array([np.linalg.inv(d) for d in dg])
    # the inverse element of the (single) group: rotation by -phi, a mirror
    # itself (in the double group dg1 may be the other preimage, -dg[invg])
    ng = len(dg)
    ops = [lm_matrix("R" if kind == "R" else "M", k * math.pi / 12.0, 1) for kind, k, _ in els]
    invg = np.zeros(ng, np.int32)
    for a in range(ng):
        inv = np.linalg.inv(ops[a])
        invg[a] = 1 + next(b for b in range(ng) if np.allclose(ops[b], inv, atol=1e-12))
    return dg, dg1, invg


def _mus(lmax):
    out = []
    for j2 in range(1, 2 * lmax + 2, 2):
        for l in ((j2 - 1) // 2, (j2 - 1) // 2 + 1):
            if l > lmax:
                continue
            out += [mu2 / 2.0 for mu2 in range(-j2, j2 + 1, 2)]
    return np.array(out)
