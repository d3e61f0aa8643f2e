"""2D k-point meshes of the reference (skkr_kpar_v6/main/kmesh2d.f:5-254):
the equidistant irreducible wedge (speck2.f), the irreducible BZ of a point
group (ibz.f) or the full BZ (fullbz.f), the rotation into the lattice's
orientation and the boundary weights, with the lattice set-up these rely on
(primkr2d.f, rgroup.f, and the holohedral group per lattice of
nrsp-scf/readgeom.f:51).  The loops and the order of operations are the
reference's, so the point sets and weights are the same numbers.

Covered: equidistant points (kunning = 0, speck2.f) on all five 2D lattices
and Cunningham special points (kunning = 1, setspk.f and speck.f; the
oblique lattice falls back to equidistant points as kmesh2d.f does); intbz
0/1 (the irreducible BZ of any allowed group, ibz.f with its skip tables)
and intbz 2 (full BZ); single k (iset = 0).
"""
from __future__ import annotations

import math

import numpy as np

BRAVAIS = {1: "oblique", 2: "centered rectangular", 3: "primitive rectangular", 4: "square", 5: "hexagonal"}
HOLOHEDRY = {1: "C2", 2: "C2v", 3: "C2v", 4: "C4v", 5: "C6v"}   # nrsp-scf/readgeom.f:51


def _rot(v, phi):
    """vecrot.f: v <- R(phi) v."""
    c, s = math.cos(phi), math.sin(phi)
    return (c * v[0] + (-s) * v[1], s * v[0] + c * v[1])


def primkr2d(ilat, idgroup, a2d, b2d=0.0, d2d=0.0):
    """main/primkr2d.f: real-space and reciprocal lattice vectors of the 2D
    Bravais lattice, the rotation arot (= phik) some group settings ask for,
    the input group name (iogroup) and the normalised one (idgroup).  beta
    and delta come back as primkr2d leaves them (it fixes them for lattices
    2-5)."""
    pi = 4.0 * math.atan(1.0)
    iogroup = idgroup
    arot = 0.0
    ax = (a2d, 0.0)
    allowed = {1: ("C1", "C2"),
               2: ("C1", "C2", "Csx", "Csy", "C2v", "C2vx"),
               3: ("C1", "C2", "Csx", "Csy", "C2v", "C2vx"),
               4: ("C1", "C2", "C4", "C4v", "C2vx", "C2vd", "Csx", "Csy", "Csd1", "Csd2"),
               5: ("C1", "C2", "C3", "C6", "C3vA", "C3vB", "C6v", "C2vx", "C2v2", "C2v3",
                   "Csx", "Csx2", "Csx3", "Csy", "Csy2", "Csy3")}
    if ilat not in allowed:
        raise ValueError("2D lattice type must be 1..5")
    if idgroup not in allowed[ilat]:
        raise ValueError(f"group {idgroup} not allowed on the {BRAVAIS[ilat]} lattice")
    if ilat == 1:
        if d2d > 0.5 or b2d >= 1:
            raise ValueError("beta, delta out of range for the oblique lattice")
        ay = (a2d * d2d, a2d * b2d)
    elif ilat in (2, 3):
        if b2d >= 1:
            raise ValueError("beta out of range for a rectangular lattice")
        d2d = 0.5 if ilat == 2 else 0.0
        ay = (a2d * d2d, a2d * d2d * b2d) if ilat == 2 else (a2d * d2d, a2d * b2d)
    elif ilat == 4:
        d2d, b2d = 0.0, 1.0
        ay = (a2d * d2d, a2d * b2d)
    else:
        b2d = 1.0
        ay = (a2d / 2.0, a2d * math.sqrt(3.0) / 2.0)
    axc, ayc = ax, ay
    rot = {"Csy": pi / 2.0}
    if ilat == 4:
        rot.update({"Csd1": pi / 4.0, "C2vd": pi / 4.0, "Csd2": -pi / 4.0})
    if ilat == 5:
        rot.update({"Csx2": pi / 3.0, "Csx3": -pi / 3.0, "Csy2": -pi / 6.0, "C2v2": -pi / 6.0,
                    "Csy3": pi / 6.0, "C2v3": pi / 6.0})
    if ilat >= 2 and idgroup in rot:
        arot = rot[idgroup]
        ax, ay = _rot(ax, arot), _rot(ay, arot)
    if idgroup.startswith("Cs"):
        idgroup = "Cs"
    elif idgroup in ("C2vx", "C2vd", "C2v2", "C2v3"):
        idgroup = "C2v"
    v1 = (ay[1], -ay[0])
    v2 = (-ax[1], ax[0])
    dnt1 = ax[0] * v1[0] + ax[1] * v1[1]
    dnt2 = ay[0] * v2[0] + ay[1] * v2[1]
    bx = (2.0 * pi * v1[0] / dnt1, 2.0 * pi * v1[1] / dnt1)
    by = (2.0 * pi * v2[0] / dnt2, 2.0 * pi * v2[1] / dnt2)
    v1 = (ayc[1], -ayc[0])
    v2 = (-axc[1], axc[0])
    dnt1 = axc[0] * v1[0] + axc[1] * v1[1]
    dnt2 = ayc[0] * v2[0] + ayc[1] * v2[1]
    bxc = (2.0 * pi * v1[0] / dnt1, 2.0 * pi * v1[1] / dnt1)
    byc = (2.0 * pi * v2[0] / dnt2, 2.0 * pi * v2[1] / dnt2)
    return dict(ilat=ilat, a2d=a2d, beta=b2d, delta=d2d, ax=ax, ay=ay, bx=bx, by=by, phik=arot,
                idgroup=idgroup, iogroup=iogroup, axc=axc, ayc=ayc, bxc=bxc, byc=byc)


def rgroup(idgroup):
    """main/rgroup.f: the 2 x 2 matrices of the point group's elements, in the
    reference's order."""
    pi = math.acos(-1.0)

    def R(t):
        return np.array([[math.cos(t), -math.sin(t)], [math.sin(t), math.cos(t)]])

    E, Mx, My = np.eye(2), np.array([[-1.0, 0.0], [0.0, 1.0]]), np.array([[1.0, 0.0], [0.0, -1.0]])

    def mirror(t):  # tripmtr(R(t), diag(1,-1), R(-t))
        return R(t) @ My @ R(-t)

    tab = {
        "C1": [E],
        "Cs": [E, My],
        "C2": [E, R(pi)],
        "C3": [E, R(2 * pi / 3), R(-2 * pi / 3)],
        "C4": [E, R(pi / 2), R(-pi / 2), R(pi)],
        "C6": [E, R(2 * pi / 6), R(-2 * pi / 6), R(2 * pi / 3), R(-2 * pi / 3), R(pi)],
        "C2v": [E, R(pi), Mx, My],
        "C3vA": [E, R(2 * pi / 3), R(-2 * pi / 3), My, mirror(-pi / 3), mirror(pi / 3)],
        "C3vB": [E, R(2 * pi / 3), R(-2 * pi / 3), Mx, mirror(pi / 6), mirror(-pi / 6)],
        "C4v": [E, R(pi / 2), R(-pi / 2), R(pi), Mx, My, np.array([[0.0, -1.0], [-1.0, 0.0]]),
                np.array([[0.0, 1.0], [1.0, 0.0]])],
        "C6v": [E, R(2 * pi / 6), R(-2 * pi / 6), R(2 * pi / 3), R(-2 * pi / 3), R(pi), Mx,
                mirror(pi / 6), mirror(-pi / 6), My, mirror(-pi / 3), mirror(pi / 3)],
    }
    # rgroup.f sets the matrices of the pure rotations element by element
    # (a(1,2) = -sin, a(2,1) = sin); R(t) above is the same matrix
    return tab[idgroup]


def speck2(n, ilat, alat, beta, delta):
    """main/speck2.f: equidistant points of the irreducible wedge of the
    lattice (its full point group), scaled by pi / alat.  The weights are
    speck2's own (kmesh2d recomputes them)."""
    xk, yk, wk = [], [], []
    if ilat == 1:
        dl = -delta if delta < 0.0 else delta
        m = int(n / math.sqrt(beta * beta + dl * dl)) + 1
        dk1 = (2.0 / n, -2.0 * dl / (n * beta))
        dk2 = (0.0, 2.0 / (m * beta))
        bod = beta / dl
        bo1md = beta / (1.0 - dl)
        tiny = 1.0e-6 * dk2[1]
        xmax = 1.0 + (1.0 / bod + 1.0 / bo1md) / (bod + bo1md)
        for i in range(1, n + 1):
            for j in range(-m + 1, m + 1):
                x = (i - 0.5) * dk1[0] + (j - 0.5) * dk2[0]
                y = (i - 0.5) * dk1[1] + (j - 0.5) * dk2[1]
                if x > xmax:
                    continue
                ymax1 = 1.0 / beta
                ymax2 = -bo1md * (x - 1.0) + 1.0 / bo1md
                ymax = min(ymax1, ymax2)
                ymin1 = -1.0 / beta
                ymin2 = bod * (x - 1.0) - 1.0 / bod
                ymin = max(ymin1, ymin2)
                dymax1, dymax2 = ymax1 - y, ymax2 - y
                dymin1, dymin2 = y - ymin1, y - ymin2
                if not (ymin <= y <= ymax):
                    continue
                if dymax1 < tiny:
                    w = 0.25 if dymax2 < tiny else 0.5
                elif dymax2 < tiny:
                    w = 0.25 if dymin1 < tiny else 0.5
                elif dymin1 < tiny:
                    w = 0.25 if dymin2 < tiny else 0.5
                elif dymin2 < tiny:
                    w = 0.5
                else:
                    w = 1.0
                xk.append(x); yk.append(y); wk.append(w)
        if delta < 0.0:
            yk = [-y for y in yk]
    elif ilat == 2:
        dkx = 2.0 / float(n + 1)
        dky = 2.0 / float(n + 2) / beta
        tiny = dkx * 1.0e-6
        x = dkx / 2.0
        while x <= 2.0:
            difx = abs(x - 2.0)
            y = dky / 2.0
            ymax = -x * beta + (beta + 1.0 / beta)
            while y <= ymax:
                dify = abs(y - ymax)
                if difx < tiny:
                    w = 0.25 if dify < tiny else 0.5
                else:
                    w = 0.5 if dify < tiny else 1.0
                xk.append(x); yk.append(y); wk.append(w)
                y = y + dky
            x = x + dkx
    elif ilat == 3:
        dkx = 1.0 / float(n)
        dky = (1.0 / beta) / float(n)
        x = dkx / 2.0
        while x <= 1.0:
            y = dky / 2.0
            while y <= 1.0 / beta:
                xk.append(x); yk.append(y); wk.append(1.0)
                y = y + dky
            x = x + dkx
    elif ilat == 4:
        for i in range(1, n + 1):
            x = 0.5 * float(2 * i - 1) / n
            for j in range(1, i + 1):
                y = 0.5 * float(2 * j - 1) / n
                xk.append(x); yk.append(y); wk.append(0.5 if i == j else 1.0)
    else:
        dxk = 2.0 / float(n)
        dyk = dxk / math.sqrt(3.0)
        for nu in (1, 2):
            anu3 = float(nu) / 3.0
            for i in range(1, n + 1):
                for j in range(i, n + 1):
                    iy = i + 2 * j + nu - 3
                    if iy > n:
                        continue
                    ajm = 1.0
                    if i == j or iy == n:
                        ajm = 2.0
                    if i == j and iy == n:
                        ajm = 6.0
                    wk.append(1.0 / ajm)
                    xk.append((float(i - 1) + anu3) * dxk)
                    yk.append(iy * dyk)
        phi = -3.1415926535897932 / 3.0
        for t in range(len(xk)):
            xk[t], yk[t] = _rot((xk[t], yk[t]), phi)
    const = 3.1415926535897932 / alat
    return [const * x for x in xk], [const * y for y in yk], wk


# setspk.f: the requested number of Cunningham points per k-set and lattice,
# and the largest k-set each lattice takes (iset is clamped to it)
_SETSPK = {1: (8, {1: 1, 2: 2, 3: 4, 4: 8, 5: 16, 6: 32, 7: 64, 8: 128}),
           2: (4, {1: 1, 2: 4, 3: 16, 4: 64}),
           3: (4, {1: 1, 2: 4, 3: 16, 4: 64}),
           4: (5, {1: 1, 2: 3, 3: 10, 4: 36, 5: 136}),
           5: (6, {1: 1, 2: 3, 3: 6, 4: 18, 5: 45, 6: 135})}
# log(2.0) and log(4.0) of speck.f are single-precision constants
_LOG2_F32 = float(np.float32(math.log(2.0)))
_LOG4_F32 = float(np.float32(math.log(4.0)))


def setspk(ilat, iset):
    """main/setspk.f: (clamped iset, requested number of special points)."""
    top, tab = _SETSPK[ilat]
    iset = min(iset, top)
    return iset, tab[iset]


def speck(nspk, lat):
    """main/speck.f:5-254: Cunningham special points of the irreducible wedge
    of the centred and primitive rectangular, square and hexagonal lattices
    (the smallest set with at least nspk points), in the unrotated reciprocal
    basis bxc, byc of primkr2d.  kmesh2d never asks for the oblique lattice."""
    stol = tol = 1.0e-6
    ilat = lat["ilat"]
    rlvax, rlvay = lat["bxc"]
    rlvby = lat["byc"][1]
    sr3 = math.sqrt(3.0)
    if ilat in (2, 3):
        xkg, ykg = rlvax / 2.0, rlvby / 8.0
        if ilat == 3:
            xkg, ykg = xkg / 2.0, ykg * 2.0
        px, py = [xkg], [ykg]
        if nspk == 1:
            return px, py
        nt = int(math.log(float(nspk)) / _LOG4_F32 + 1.0 - stol)
        nspk = 4 ** (nt + 1) // 4
        while True:
            xkg, ykg = xkg / 2.0, ykg / 2.0
            nx, ny = [], []
            for ox, oy in zip(px, py):
                x1, y1, x2, y2 = ox + xkg, oy + ykg, ox - xkg, oy - ykg
                nx += [x1, x1, x2, x2]
                ny += [y1, y2, y1, y2]
            px, py = nx, ny
            if len(px) >= nspk:
                break
        if ilat == 2:
            beta = -rlvax / rlvay
            for i in range(len(px)):
                t = py[i] + 0.5 * px[i] / beta - 0.75 * rlvax / beta
                if t > 0.0:
                    px[i], py[i] = -px[i] + rlvax, -py[i] - rlvay
        return px, py
    if ilat == 4:
        xkg = ykg = rlvax / 4.0
        px, py = [xkg], [ykg]
        if nspk <= 1:
            return px, py
        nt = int(math.log((-1.0 + math.sqrt(1.0 + 8.0 * float(nspk))) / 2.0) / _LOG2_F32 + 1.0 - stol)
        nu = 2 ** (nt + 1) // 2
        nspk = (nu * nu + nu) // 2
        insp = 0
        while True:
            nt = 2 ** (insp + 1) // 2
            if (nt * nt + nt) // 2 >= nspk:
                return px, py
            xkg, ykg = xkg / 2.0, ykg / 2.0
            nx, ny = [], []
            for ox, oy in zip(px, py):
                xs, ys = (ox + xkg, ox - xkg), (oy + ykg, oy - ykg)
                for xv in xs:
                    for yv in ys:
                        if xv < yv:
                            continue
                        nx.append(xv)
                        ny.append(yv)
            px, py = nx, ny
            insp += 1
    if ilat == 5:
        hx = (1.0, 0.5, -0.5, -1.0, -0.5, 0.5)
        hy = (0.0, 0.5 * sr3, 0.5 * sr3, 0.0, -0.5 * sr3, -0.5 * sr3)
        sa = 0.5 * rlvax
        xs = [4.0 * sa / 3.0, 2.0 * sa / 3.0]
        ys = [0.0, 2.0 * sa / 3.0 / sr3]
        px, py = [xs[1]], [ys[1]]
        if nspk <= 1:
            return px, py
        ix = 2
        while True:
            ix = 2 - (ix + 1) % 2
            xs[ix - 1] /= 3.0
            ys[ix - 1] /= 3.0
            xkg, ykg = xs[ix - 1], ys[ix - 1]
            nx, ny = [], []
            for xa, ya in zip(px, py):
                for j in range(1, 7):
                    jm = j % 6
                    xap = hx[jm] * xkg + hy[jm] * ykg + xa
                    yap = -hy[jm] * xkg + hx[jm] * ykg + ya
                    if yap < -tol or (yap - xap / sr3) > tol or (yap + xap * sr3 - 2.0 / sr3 * rlvax) > tol:
                        continue
                    if any(abs(xap - qx) < tol and abs(yap - qy) < tol for qx, qy in zip(nx, ny)):
                        continue
                    nx.append(xap)
                    ny.append(yap)
            px, py = nx, ny
            if len(px) >= nspk:
                return px, py
    raise ValueError("speck: lattice type must be 2..5")


def _ibz(lat, kunning, xk1, yk1):
    """main/ibz.f: the wedge of the lattice's holohedral group (xk1, yk1)
    unfolded into the irreducible BZ of the input group, with the skip
    tables that keep every edge point once."""
    tol = 1.0e-8
    pi = 4.0 * math.atan(1.0)
    ilat, idgroup, iogroup = lat["ilat"], lat["idgroup"], lat["iogroup"]
    amat = rgroup(HOLOHEDRY[ilat])
    ng = len(amat)
    xo, yo = [], []
    for x, y in zip(xk1, yk1):
        ngeff = ng
        skip = set()
        if ilat == 1:
            if idgroup == "C2":
                ngeff = 1
            if (abs(x - 0.0) < tol) if kunning == 0 else (abs(y - 0.0) < tol):
                ngeff = 1
        elif ilat in (2, 3):
            if idgroup == "C2v":
                ngeff = 1
            elif idgroup in ("Cs", "C2"):
                if iogroup == "Csy":
                    if abs(y - 0.0) < tol:
                        ngeff = 1
                    else:
                        skip, ngeff = {2}, 3
                elif abs(x - 0.0) < tol:
                    ngeff = 1
                else:
                    skip = {2, 3}
            elif abs(x - 0.0) < tol or abs(y - 0.0) < tol:
                ngeff = 2
        elif ilat == 4:
            if idgroup == "C4v":
                ngeff = 1
            elif idgroup in ("C2v", "C4"):
                if iogroup == "C2vd":
                    ngeff, skip = 6, {2, 3, 4, 5}
                    if abs(y - 0.0) < tol:
                        ngeff = 1
                elif abs(x - y) < tol:
                    ngeff = 1
                else:
                    skip = set(range(2, ng))
            elif idgroup in ("C2", "Cs"):
                skip = {3, 4, 7}
                if iogroup in ("Csd1", "Csd2"):
                    skip.add(5)
                    if abs(x - y) < tol:
                        ngeff = 6
                    elif abs(y - 0.0) < tol:
                        ngeff = 2
                else:
                    skip.add(6)
                    if abs(x - y) < tol:
                        ngeff = 2
                    elif abs(y - 0.0) < tol:
                        ngeff = 5
            elif abs(x - y) < tol or abs(y - 0.0) < tol:
                ngeff = 4
        else:
            phi = _asin_phi(x, y)
            at0, at30 = abs(phi - 0.0) < tol, abs(phi - pi / 6.0) < tol
            if idgroup == "C6v":
                ngeff = 1
            elif idgroup in ("C6", "C3vB"):
                if at0:
                    ngeff = 1
                else:
                    ngeff, skip = 10, set(range(2, 10))
            elif idgroup == "C3vA":
                if at30:
                    ngeff = 1
                else:
                    ngeff, skip = 8, set(range(2, 8))
            elif idgroup == "C3":
                if at30:
                    ngeff = 2
                else:
                    skip = set(range(3, ngeff)) - {8}
                    if at0:
                        skip.add(8)
            elif idgroup == "C2v":
                if iogroup in ("C2v2", "C2v3"):
                    ngeff, skip = 10, set(range(2, 8)) | {9}
                    if at30:
                        skip.add(8)
                    if at0:
                        ngeff = 8
                elif at30 or at0:
                    ngeff = 2
                else:
                    ngeff, skip = 8, set(range(3, 8))
            elif idgroup in ("Cs", "C2"):
                if iogroup in ("Csy", "Csy2", "Csy3"):
                    skip, ngeff = set(range(4, 8)), 10
                    if at30:
                        skip |= {8, 10}
                    if at0:
                        ngeff = 3
                else:
                    skip = set(range(3, ngeff)) - {4, 7, 8}
                    if at30:
                        ngeff = 4
                    if at0:
                        ngeff = 7
            elif at30 or at0:
                ngeff = 6
        for ig in range(1, ngeff + 1):
            if ig not in skip:
                a = amat[ig - 1]
                xo.append(a[0, 0] * x + a[0, 1] * y)
                yo.append(a[1, 0] * x + a[1, 1] * y)
    return xo, yo


def _findangles_phi(x, y):
    """main/findangles.f with z = 0: the azimuth in [0, 2 pi)."""
    tol = 1.0e-5
    pi = math.acos(-1.0)
    dx, dy = abs(x), abs(y)
    rxy = math.sqrt(x * x + y * y)
    if rxy < tol or (dx < tol and dy < tol):
        return 0.0
    if dy < tol:
        return 0.0 if x > 0.0 else (pi if x < 0.0 else 0.0)
    if dx < tol:
        return 0.5 * pi if y > 0.0 else (1.5 * pi if y < 0.0 else 0.0)
    phi = math.acos(dx / rxy)
    if x < 0.0 and y > 0.0:
        phi = pi - phi
    if x < 0.0 and y < 0.0:
        phi = pi + phi
    if x > 0.0 and y < 0.0:
        phi = 2.0 * pi - phi
    return phi


def _asin_phi(x, y):
    r = math.sqrt(x * x + y * y)
    return math.asin(y / r) if r > 0.0 else float("nan")


def _fullbz(lat, iset, kunning, xk1, yk1):
    """main/fullbz.f: the wedge unfolded with the holohedral group (iset > 0)
    or the images of one k under the input group (iset = 0)."""
    tol = 1.0e-8
    pi = 4.0 * math.atan(1.0)
    ilat = lat["ilat"]
    if iset == 0:
        v = (xk1[0], yk1[0])
        if abs(lat["phik"] - 0.0) > tol:
            v = _rot(v, lat["phik"])
        pts = [(a[0, 0] * v[0] + a[0, 1] * v[1], a[1, 0] * v[0] + a[1, 1] * v[1])
               for a in rgroup(lat["idgroup"])]
        keep = [True] * len(pts)
        for i in range(len(pts) - 1):
            for j in range(i + 1, len(pts)):
                if math.sqrt((pts[i][0] - pts[j][0]) ** 2 + (pts[i][1] - pts[j][1]) ** 2) - 0.0 < tol:
                    keep[j] = False
        pts = [p for p, k in zip(pts, keep) if k]
        return [p[0] for p in pts], [p[1] for p in pts]
    amat = rgroup(HOLOHEDRY[ilat])
    xo, yo = [], []
    for x, y in zip(xk1, yk1):
        ngeff = len(amat)
        if ilat == 1:
            if (abs(x - 0.0) < tol) if kunning == 0 else (abs(y - 0.0) < tol):
                ngeff = 1
        elif ilat in (2, 3):
            if abs(y - 0.0) < tol or abs(x - 0.0) < tol:
                ngeff = 2
        elif ilat == 4:
            if abs(x - 0.0) < tol or abs(x - y) < tol:
                ngeff = 4
        else:
            phi = _asin_phi(x, y)
            if abs(phi - 0.0) < tol or abs(phi - pi / 6.0) < tol:
                ngeff = 6
        for a in amat[:ngeff]:
            xo.append(a[0, 0] * x + a[0, 1] * y)
            yo.append(a[1, 0] * x + a[1, 1] * y)
    return xo, yo


def kmesh2d(intbz, iset, lat, park=(0.0, 0.0), kunning=0):
    """main/kmesh2d.f:5-254: k-points (nk, 2) and normalised weights (nk,) of
    the mesh iset of the lattice set up by primkr2d (lat)."""
    tol = 1.0e-10
    pi = math.acos(-1.0)
    ilat, idgroup, iogroup, phik = lat["ilat"], lat["idgroup"], lat["iogroup"], lat["phik"]
    if iset == 0:
        xk0, yk0 = [park[0]], [park[1]]
    else:
        if ilat == 1 and kunning == 1:
            kunning = 0   # kmesh2d.f: no Cunningham points on the oblique lattice
        if kunning == 0:
            xk0, yk0, _ = speck2(iset, ilat, lat["a2d"], lat["beta"], lat["delta"])
        else:
            iset, nspk = setspk(ilat, iset)
            xk0, yk0 = speck(nspk, lat)
    if intbz <= 1:
        if iset == 0:
            xk, yk = list(xk0), list(yk0)
        else:
            xk, yk = _ibz(lat, kunning, xk0, yk0)
    else:
        xk, yk = _fullbz(lat, iset, kunning, xk0, yk0)
    phi = phik
    if intbz <= 1 and iset != 0:
        if ilat == 4:
            if iogroup == "Csy":
                phi = 0.0
            if iogroup == "Csd2":
                phi = -phik
        elif ilat == 5:
            if iogroup in ("Csx2", "Csx3"):
                phi = 0.0
            if iogroup in ("Csy2", "Csy3"):
                phi = pi / 2.0
            if iogroup == "C2v2":
                phi = -phik
    if intbz <= 1 or iset != 0:
        if abs(phi - 0.0) > tol:
            for t in range(len(xk)):
                xk[t], yk[t] = _rot((xk[t], yk[t]), phi)
    bx, by = lat["bx"], lat["by"]
    wk = []
    for x, y in zip(xk, yk):
        ibound = 0
        for i in range(-2, 3):
            for j in range(-2, 3):
                if i == 0 and j == 0:
                    continue
                gx = float(i) * bx[0] + float(j) * by[0]
                gy = float(i) * bx[1] + float(j) * by[1]
                g2o2 = 0.5 * (gx * gx + gy * gy)
                gk = gx * x + gy * y
                if abs(gk - g2o2) < tol:
                    ibound += 1
        if ibound == 0:
            w = 1.0
        elif ibound == 1:
            w = 0.5
        elif ibound == 2:
            w = 0.25 if ilat in (3, 4) else 1.0 / 3.0
        else:
            raise RuntimeError("a k-point on more than two BZ boundaries")
        wk.append(w)
    if intbz <= 1 and iset != 0:
        for t, (x, y) in enumerate(zip(xk, yk)):
            ph = _findangles_phi(x, y)
            ax0, ay0, axy = abs(x - 0.0) < tol, abs(y - 0.0) < tol, abs(x - y) < tol
            if idgroup in ("Cs", "C2"):
                if ilat == 1 and kunning == 0:
                    if ax0:
                        wk[t] *= 0.5
                elif ay0:
                    wk[t] *= 0.5
            elif idgroup in ("C4", "C2v"):
                if ax0:
                    wk[t] *= 0.25 if ay0 else 0.5
                elif ay0:
                    wk[t] *= 0.5
            elif idgroup == "C4v":
                if ax0:
                    wk[t] *= 0.125 if ay0 else 0.5
                elif axy:
                    wk[t] *= 0.5
            elif idgroup == "C3":
                if ay0:
                    wk[t] = wk[t] / 3.0 if ax0 else 0.5 * wk[t]
                elif abs(ph - 2.0 * pi / 3.0) < tol:
                    wk[t] *= 0.5
            elif idgroup == "C3vA":
                if ay0:
                    wk[t] = wk[t] / 6.0 if ax0 else 0.5 * wk[t]
                elif abs(ph - pi / 3.0) < tol:
                    wk[t] *= 0.5
            elif idgroup in ("C3vB", "C6"):
                if ax0 and ay0:
                    wk[t] = wk[t] / 6.0
                elif abs(ph - pi / 6.0) < tol or abs(ph + pi / 6.0) < tol:
                    wk[t] *= 0.5
            elif idgroup == "C6v":
                if ax0:
                    wk[t] = wk[t] / 12.0 if ay0 else 0.5 * wk[t]
                elif abs(ph - pi / 6.0) < tol:
                    wk[t] *= 0.5
    wsum = 0.0
    for w in wk:
        wsum += w
    return np.array([xk, yk]).T.reshape(-1, 2), np.array([w / wsum for w in wk])
