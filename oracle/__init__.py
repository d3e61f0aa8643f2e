"""CPU FP64 oracle of the per-(E,k) layer kernel -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use this package, and only as the checker.  The product
path (``skkr_budapest_amd``) never imports it.

``liboracle.so`` is built from ``skkr_oracle.c`` by ``make -C oracle``; each C
routine cites the reference subroutine it restates.  This module wraps it with
numpy arrays indexed ``A[row, col]``.

Parity status of each routine (what pins it):
  gjinv, surfgfd, dugo, tau2d, packer, k-loop/BZ sum -- golden vectors from
      the reference's own Fortran compiled from /root/reference (kpar and
      dbogo twins, tests/golden), plus dense-inverse and closed-form checks.
  fourier -- no reference routine; brute force, k=0, Hermiticity,
      reciprocal-lattice periodicity and linearity.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

STATUS = {0: "ok", 1: "singular pivot", 2: "decimation diverged",
          3: "decimation itermax exceeded", 4: "out of memory"}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i, d = ctypes.c_int, ctypes.c_double
        L.orc_gjinv.argtypes = [P, i, i]
        L.orc_surfgfd.argtypes = [P, P, P, P, i, i, i, d, P]
        L.orc_dugo.argtypes = [P, P, P, P, i, i, ctypes.c_char, i, d, P, P]
        L.orc_tau2d.argtypes = [P, P, P, P, i, i, i, P, P, d, P]
        L.orc_tau2d_pl.argtypes = [P, P, P, ctypes.c_size_t, P, i, i, i, P, P, d, P]
        L.orc_fourier.argtypes = [i, P, P, i, P, P]
        L.orc_pl_blocks.argtypes = [P, P, i, P, P, P]
        L.orc_tau_bz.argtypes = [i, i, i, i, P, P, P, P, P, P, i, d, i, P, P, i,
                                 P, P, P, P, P]
        L.orc_phystau.argtypes = [i, P, P, P, P, i, i]
        L.orc_phystau.restype = None
        for f in (L.orc_gjinv, L.orc_surfgfd, L.orc_dugo, L.orc_tau2d, L.orc_tau2d_pl, L.orc_tau_bz):
            f.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _cm(a):
    """[..., row, col] -> column-major complex128 buffer."""
    return np.ascontiguousarray(np.swapaxes(np.asarray(a, np.complex128), -1, -2))


def _uncm(b):
    return np.ascontiguousarray(np.swapaxes(b, -1, -2))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def gjinv(a):
    b = _cm(a)
    st = lib().orc_gjinv(_p(b), a.shape[0], a.shape[0])
    return _uncm(b), st


def phystau(tau, tinv1, tinv2, a, idiag=1):
    """Screened -> physical tau of one site (phystau.f); a: per-index
    screening parameters (n,).  Returns the new (n, n) matrix."""
    n = tau.shape[0]
    t = _cm(tau)
    t1, t2 = _cm(tinv1), _cm(tinv2)
    av = np.ascontiguousarray(a, np.complex128)
    lib().orc_phystau(n, _p(av), _p(t1), _p(t2), _p(t), int(idiag), n)
    return _uncm(t)


def phys_traces(tau, tinv, a, l_of_index, nl):
    """Per-l traces of phystau(tau) (cpacoord.f:300-309 taukdiag)."""
    tp = phystau(tau, tinv, tinv, a, 1)
    d = np.diagonal(tp)
    out = np.zeros(nl, np.complex128)
    for i, l in enumerate(l_of_index):
        out[l] += d[i]
    return out


def surfgfd(ml, m0, mr, itermax=50, errmax=1e-8):
    n = m0.shape[0]
    bl, b0, br = _cm(ml), _cm(m0), _cm(mr)
    x = np.zeros_like(b0)
    it = ctypes.c_int(0)
    st = lib().orc_surfgfd(_p(bl), _p(b0), _p(br), _p(x), n, n, itermax, errmax,
                           ctypes.byref(it))
    return _uncm(x), it.value, st


def pl_blocks(g_intra_k, g_inter_k):
    n = g_intra_k.shape[0]
    gi, go = _cm(g_intra_k), _cm(g_inter_k)
    m00, m01, m10 = (np.zeros_like(gi) for _ in range(3))
    lib().orc_pl_blocks(_p(gi), _p(go), n, _p(m00), _p(m01), _p(m10))
    return _uncm(m00), _uncm(m01), _uncm(m10)


def dugo(m00g, m01, m10, tinv_lead, side, itermax=50, errmax=1e-8):
    s, lk = tinv_lead.shape[0], tinv_lead.shape[1]
    a, b, c, t = _cm(m00g), _cm(m01), _cm(m10), _cm(tinv_lead)
    xy = np.zeros_like(a)
    it = ctypes.c_int(0)
    st = lib().orc_dugo(_p(a), _p(b), _p(c), _p(t), lk, s, side.encode(), itermax,
                        errmax, _p(xy), ctypes.byref(it))
    return _uncm(xy), it.value, st


def tau2d(m00g, m01, m10, tinv, P, x, y, w=1.0):
    lk = tinv.shape[1]
    s = tinv.shape[0] // P
    a, b, c, t = _cm(m00g), _cm(m01), _cm(m10), _cm(tinv)
    xb, yb = _cm(x), _cm(y)
    tau = np.zeros((P * s, lk, lk), np.complex128)
    st = lib().orc_tau2d(_p(a), _p(b), _p(c), _p(t), lk, s, P, _p(xb), _p(yb),
                         float(w), _p(tau))
    return _uncm(tau), st


def symmetrize(tau_bz, dg, dg1):
    """The point-group sum of nrsp-scf/cpacoord.f:294-299 (tripmt, addmatc):
    (1/ng) sum_g dg_g tau dg1_g per site, here applied to the BZ sum, which
    the per-k form equals by linearity."""
    tau_bz = np.asarray(tau_bz, np.complex128)
    out = np.zeros_like(tau_bz)
    for d, d1 in zip(dg, dg1):
        out += d[None] @ tau_bz @ d1[None]
    return out / len(dg)


def tau2d_layered(m00g, m01, m10, tinv, x, y, w=1.0):
    """tau2d.f:2-152 with bulkgeo false: PL p reads its own packer blocks
    m00g[p], m01[p], m10[p] (plane ipl = p + 1, :55-56,90-91)."""
    P, n = m00g.shape[0], m00g.shape[1]
    lk = tinv.shape[1]
    s = tinv.shape[0] // P
    a, b, c, t = _cm(m00g), _cm(m01), _cm(m10), _cm(tinv)
    xb, yb = _cm(x), _cm(y)
    tau = np.zeros((P * s, lk, lk), np.complex128)
    st = lib().orc_tau2d_pl(_p(a), _p(b), _p(c), n * n, _p(t), lk, s, P, _p(xb), _p(yb),
                            float(w), _p(tau))
    return _uncm(tau), st


def fourier(R, GR, k):
    n = GR.shape[1]
    Rb = np.ascontiguousarray(R, np.float64)
    gb = _cm(GR)
    kb = np.ascontiguousarray(k, np.float64)
    out = np.zeros((n, n), np.complex128)
    lib().orc_fourier(len(Rb), _p(Rb), _p(gb), n, _p(kb), _p(out))
    return _uncm(out)


def tau_bz(system, E, k, w, nworkers=1, want_tau_k=False, itermax=None,
           errmax=None, tinv_int=None):
    """The whole k-loop of one energy.  Returns dict with tau_bz (sites,lk,lk),
    trk (nk, sites) traces of the unweighted per-k tau, iters (nk, 2) and
    status (nk,), plus tau_k (nk, sites, lk, lk) when asked.  tinv_int
    replaces the interface t^-1 of the system (the CPA's coherent one)."""
    cfg = system.cfg
    lk, s, P = cfg.lk, cfg.s, cfg.P
    tint, tl, tr = system.tinv(E)
    if tinv_int is not None:
        tint = np.asarray(tinv_int, np.complex128)
    R = np.ascontiguousarray(system.R, np.float64)
    gi, go = _cm(system.G_intra), _cm(system.G_inter)
    ti, tlb, trb = _cm(tint), _cm(tl), _cm(tr)
    k = np.ascontiguousarray(k, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    nk = len(w)
    ns = P * s
    tau = np.zeros((ns, lk, lk), np.complex128)
    tau_k = np.zeros((nk, ns, lk, lk), np.complex128) if want_tau_k else None
    trk = np.zeros((nk, ns), np.complex128)
    iters = np.zeros((nk, 2), np.int32)
    status = np.zeros(nk, np.int32)
    st = lib().orc_tau_bz(lk, s, P, len(R), _p(R), _p(gi), _p(go), _p(ti), _p(tlb),
                          _p(trb), int(itermax or cfg.itermax),
                          float(errmax or cfg.errmax), nk, _p(k), _p(w),
                          int(nworkers), _p(tau),
                          _p(tau_k) if tau_k is not None else None, _p(trk),
                          _p(iters), _p(status))
    out = {"tau_bz": _uncm(tau), "trk": trk, "iters": iters, "status": status,
           "rc": st}
    if tau_k is not None:
        out["tau_k"] = _uncm(tau_k)
    return out


# ---------------------------------------------------------------------------
# Layer structure constants (kkrstr_oracle.c): Kambe Ewald sum, direct
# reciprocal sum and screening of skkr_kpar_v6/main/gstore.f.

def _kkr_lib():
    L = lib()
    if not hasattr(L, "_kkr_ready"):
        P = ctypes.c_void_p
        i, d = ctypes.c_int, ctypes.c_double
        L.orc_cerf_v.argtypes = [i, P, P]
        L.orc_incgamma_p.argtypes = [i, P, P]
        L.orc_delkam_p.argtypes = [i, P, P, P]
        L.orc_ilkam_p.argtypes = [i, P, d, P, P]
        L.orc_lattice_new.argtypes = [P, P, i]
        L.orc_lattice_new.restype = P
        L.orc_lattice_info.argtypes = [P, P, P]
        L.orc_lattice_vectors.argtypes = [P, P, P]
        L.orc_lattice_free.argtypes = [P]
        L.orc_kamstr_p.argtypes = [P, i, P, d, d, P, d, P, P]
        L.orc_directg_p.argtypes = [P, i, d, d, P, P, P, P, P]
        L.orc_gstore_p.argtypes = [P, i, i, i, P, P, d, d, P, d, P]
        L.orc_gaunt.argtypes = [i, P]
        L.orc_ylmc.argtypes = [d, d, d, i, P]
        for f in (L.orc_kamstr_p, L.orc_directg_p, L.orc_gstore_p):
            f.restype = ctypes.c_int
        L._kkr_ready = True
    return L


def cerf(z):
    """Faddeeva w(z) as cerf.f computes it (its series, thresholds and
    single-precision constants)."""
    z = np.ascontiguousarray(np.atleast_1d(z), np.complex128)
    w = np.empty_like(z)
    _kkr_lib().orc_cerf_v(len(z), _p(z), _p(w))
    return w


def incgamma(n, z):
    g = np.empty(n + 1, np.complex128)
    zz = np.array([z], np.complex128)
    _kkr_lib().orc_incgamma_p(n, _p(zz), _p(g))
    return g


def delkam(n, x, zz):
    out = np.empty(max(n + 1, 2), np.complex128)
    a, b = np.array([x], np.complex128), np.array([zz], np.complex128)
    _kkr_lib().orc_delkam_p(n, _p(a), _p(b), _p(out))
    return out[:n + 1]


def ilkam(l, alpha, t, kappa):
    out = np.empty(l + 1, np.complex128)
    a, k = np.array([alpha], np.complex128), np.array([kappa], np.complex128)
    _kkr_lib().orc_ilkam_p(l, _p(a), float(t), _p(k), _p(out))
    return out


def gaunt_table(lmax):
    """gauntcalc.f: gau over (l1,m1,l2,m2,l,m), l <= 2 lmax."""
    nl2, lg = (lmax + 1) ** 2, (2 * lmax + 1) ** 2
    g = np.empty(nl2 * nl2 * lg)
    _kkr_lib().orc_gaunt(lmax, _p(g))
    return g.reshape(nl2, nl2, lg)


def ylmc(v, lmax):
    y = np.empty((lmax + 1) ** 2, np.complex128)
    _kkr_lib().orc_ylmc(float(v[0]), float(v[1]), float(v[2]), lmax, _p(y))
    return y


class Lattice:
    """2D direct and reciprocal lattice of vector2d.f (mdir vectors each),
    the reciprocal vectors sorted into shells around the origin."""

    def __init__(self, a1, a2, mdir=20000):
        a = np.array(a1, np.float64)
        b = np.array(a2, np.float64)
        self._L = _kkr_lib().orc_lattice_new(_p(a), _p(b), mdir)
        d = np.empty(8)
        n = np.empty(3, np.int32)
        _kkr_lib().orc_lattice_info(self._L, _p(d), _p(n))
        self.distr, self.distk, self.vol, self.volbz = d[:4]
        self.b1, self.b2 = d[4:6].copy(), d[6:8].copy()
        self.maxr, self.maxk, self.maxshk = (int(x) for x in n)

    def vectors(self):
        r = np.empty((self.maxr, 2))
        g = np.empty((self.maxk, 2))
        _kkr_lib().orc_lattice_vectors(self._L, _p(r), _p(g))
        return r, g

    def __del__(self):
        try:
            _kkr_lib().orc_lattice_free(self._L)
        except Exception:
            pass

    def kamstr(self, lmax, e, k, sep, eta):
        """Kambe structure constant of one layer pair (sep = 3-vector);
        returns (alm [lm, lm'], (G shells, R shells), status)."""
        nl2 = (lmax + 1) ** 2
        alm = np.zeros(nl2 * nl2, np.complex128)
        ts = np.array([np.linalg.norm(sep), *sep], np.float64)
        ee = np.array([e], np.complex128)
        nsh = np.zeros(2, np.int32)
        st = _kkr_lib().orc_kamstr_p(self._L, lmax, _p(ee), float(k[0]), float(k[1]), _p(ts),
                                      float(eta), _p(alm), _p(nsh))
        return alm.reshape(nl2, nl2).T, tuple(int(x) for x in nsh), st

    def directg(self, lmax, e, k, sep):
        nl2 = (lmax + 1) ** 2
        gp = np.zeros(nl2 * nl2, np.complex128)
        gm = np.zeros(nl2 * nl2, np.complex128)
        sp = np.array(sep, np.float64)
        ee = np.array([e], np.complex128)
        nsh = np.zeros(1, np.int32)
        st = _kkr_lib().orc_directg_p(self._L, lmax, float(k[0]), float(k[1]), _p(sp), _p(ee),
                                       _p(gp), _p(gm), _p(nsh))
        return gp.reshape(nl2, nl2).T, gm.reshape(nl2, nl2).T, int(nsh[0]), st

    def gstore(self, lmax, s, nextra, cvec_pl, alpha, e, k, eta):
        """g2d of gstore.f (bulkgeo) for one k: array [icpl, icount, jj+s, n, m]
        (the reverse of g2d(m, n, jj, icount, icpl)); alpha[layer, l]."""
        nl2 = (lmax + 1) ** 2
        cv = np.ascontiguousarray(cvec_pl, np.float64)
        al = np.ascontiguousarray(alpha, np.complex128)
        assert cv.shape == (s + 1, 3) and al.shape == (s, lmax + 1)
        out = np.zeros(nl2 * nl2 * (2 * s + 1) * s * 3, np.complex128)
        ee = np.array([e], np.complex128)
        st = _kkr_lib().orc_gstore_p(self._L, lmax, s, nextra, _p(cv), _p(al), float(k[0]),
                                      float(k[1]), _p(ee), float(eta), _p(out))
        if st:
            raise RuntimeError(f"orc_gstore status {st}")
        return out.reshape(3, s, 2 * s + 1, nl2, nl2)


def kkr_blocks_rel(g2d, s, lmax, fac):
    """kkr_blocks in the kappa-mu basis: every lm block through relmtrx
    (packer.f with irel = 1)."""
    nl2 = (lmax + 1) ** 2
    lk = 2 * nl2
    g = np.zeros((3, s, 2 * s + 1, lk, lk), np.complex128)
    for p in range(3):
        for i in range(s):
            for jj in range(2 * s + 1):
                # relmtrx of the [m, n] block, stored back as [n, m] like g2d
                g[p, i, jj] = relmtrx(g2d[p, i, jj].T, fac, lmax).T
    return kkr_blocks(g, s, lk)


def kkr_blocks(g2d, s, nl2):
    """packer.f:36-97 (equal ninprc, irel = 0) on the g2d planes of one k
    (array [icpl, icount, jj+s, n, m] as Lattice.gstore returns it): M00, M01,
    M10 of plane 1, M01 of plane 0 and M10 of plane 2 (the outer factors of
    the lead self-energies, dugo.f:71-75)."""
    n = s * nl2

    def blk(p, i, jj):
        return g2d[p, i, jj + s].T  # g2d(m, n, jj, i, p) as [m, n]

    def plane(p):
        m00, m01, m10 = (np.zeros((n, n), np.complex128) for _ in range(3))
        for i in range(s):
            r = slice(i * nl2, (i + 1) * nl2)
            for jp in range(s):
                c = slice(jp * nl2, (jp + 1) * nl2)
                m00[r, c] = -blk(p, i, jp - i)
                if jp <= i:
                    m01[r, c] = -blk(p, i, jp + s - i)
                if jp >= i:
                    m10[r, c] = -blk(p, i, jp - s - i)
        return m00, m01, m10

    m00, m01, m10 = plane(1)
    return m00, m01, m10, plane(0)[1], plane(2)[2]


def kkr_tau_bz(lattice, lmax, s, nextra, cvec, alpha, E, eta, tinv, tinv_l, tinv_r, ks, ws,
               itermax=50, errmax=1e-8, rel=0):
    """k-loop of nrsp-scf/cpacoord.f:147-311 with bulkgeo on the oracle:
    gstore (side 'L') -> packer -> dugo('L'), dugo('R') with the outer factors
    of planes 0 and 2 -> tau2d; BZ sum of the weighted site-diagonal tau.
    tinv: (P*s, nl2, nl2); tinv_l, tinv_r: (s, nl2, nl2).  Returns
    (tau_bz [sites, nl2, nl2], per-k decimation iterations [k, 2])."""
    nl2 = (lmax + 1) ** 2
    lk = 2 * nl2 if rel else nl2
    P = tinv.shape[0] // s
    acc = np.zeros((P * s, lk, lk), np.complex128)
    its = []
    E = complex(E)
    psq = E + E * E / 274.072 ** 2 if rel == 1 else E
    for k, w in zip(ks, ws):
        g2d = lattice.gstore(lmax, s, nextra, cvec, alpha, psq, k, eta)
        if rel:
            m00, m01, m10, mtl, mtr = kkr_blocks_rel(g2d, s, lmax, psq / E)
        else:
            m00, m01, m10, mtl, mtr = kkr_blocks(g2d, s, nl2)
        tl = m00.copy()
        tr = m00.copy()
        for i in range(s):
            sl = slice(i * lk, (i + 1) * lk)
            tl[sl, sl] += tinv_l[i]
            tr[sl, sl] += tinv_r[i]
        gl, itl, st1 = surfgfd(m10, tl, m01, itermax, errmax)
        gr, itr, st2 = surfgfd(m01, tr, m10, itermax, errmax)
        if st1 or st2:
            raise RuntimeError(f"surfgfd status {st1}, {st2} at k={k}")
        y = m10 @ gl @ mtl
        x = m01 @ gr @ mtr
        tau, st = tau2d(m00, m01, m10, tinv, P, x, y, w)
        if st:
            raise RuntimeError(f"tau2d status {st} at k={k}")
        acc += tau
        its.append((itl, itr))
    return acc, np.array(its)


# ---------------------------------------------------------------------------
# Coherent-potential solver (cpa_oracle.c): tata, cpasol, cpacor of
# skkr_kpar_v6/nrsp-scf/fullcpa.f.

def _cpa_lib():
    L = lib()
    if not hasattr(L, "_cpa_ready"):
        P = ctypes.c_void_p
        i, d = ctypes.c_int, ctypes.c_double
        L.orc_tata.argtypes = [d, i, P, P, P]
        L.orc_cpasol.argtypes = [d, i, P, P, P, P, P, P, P]
        L.orc_cpacor.argtypes = [i, i, P, P, P, P, P, P, P, P]
        for f in (L.orc_tata, L.orc_cpasol, L.orc_cpacor):
            f.restype = ctypes.c_int
        L._cpa_ready = True
    return L


def tata(conc, ma, mb):
    """Initial coherent t^-1 of one site: mb (c mb + (1-c) ma)^-1 ma."""
    n = ma.shape[0]
    mc = np.zeros((n, n), np.complex128)
    a, b, c = _cm(ma), _cm(mb), _cm(mc)
    _cpa_lib().orc_tata(float(conc), n, _p(c), _p(a), _p(b))
    return _uncm(c)


def cpacor(conc, mc, ma, mb, tauc):
    """One CPA update over the sites (arrays [site, n, n]): returns the new
    coherent t^-1, tau_A, tau_B and the error (largest error3)."""
    nsite, n = ma.shape[0], ma.shape[1]
    c = _cm(mc)
    a, b, t = _cm(ma), _cm(mb), _cm(tauc)
    ta, tb = np.zeros_like(a), np.zeros_like(a)
    cc = np.ascontiguousarray(conc, np.float64)
    err = ctypes.c_double(0.0)
    st = _cpa_lib().orc_cpacor(nsite, n, _p(cc), _p(c), _p(a), _p(b), _p(t), _p(ta), _p(tb),
                               ctypes.byref(err))
    if st:
        raise RuntimeError("orc_cpacor: singular matrix")
    return _uncm(c), _uncm(ta), _uncm(tb), err.value


def cpa_bz(system, E, k, w, conc, tinv_a, tinv_b, tinv_c=None, itcpam=50, cpatol=1e-10, nworkers=1):
    """CPA loop of one energy (nrsp-scf/cpacoord.f:107-345): interface t^-1 =
    coherent mc, k-loop -> tau_c, cpacor; until error <= cpatol or itcpam
    passes.  tinv_c None starts from tata.  Returns tau_c, tau_a, tau_b,
    tinv_c (the updated mc), passes and error."""
    if tinv_c is None:
        mc = np.array([tata(c, a, b) for c, a, b in zip(conc, tinv_a, tinv_b)])
    else:
        mc = np.array(tinv_c, np.complex128)
    itcpa = 1
    while True:
        r = tau_bz(system, E, k, w, nworkers=nworkers, tinv_int=mc)
        if r["rc"]:
            raise RuntimeError(f"oracle k-loop status {r['rc']}")
        tauc = r["tau_bz"]
        mc, ta, tb, err = cpacor(conc, mc, tinv_a, tinv_b, tauc)
        concpa, cpacon = err > cpatol, itcpa < itcpam
        itcpa += 1
        if not (concpa and cpacon):
            break
    return {"tau_c": tauc, "tau_a": ta, "tau_b": tb, "tinv_c": mc, "itcpa": itcpa - 1, "error": err,
            "iters": r["iters"]}


def full_tau_k(system, E, k, itermax=None, errmax=None):
    """tau between every pair of interface sites at one k, [li, lj, lk, lk]:
    the dense inverse (LAPACK through numpy) of the finite stack of P
    principal layers with the lead self-energies of dugo on the outer PLs --
    the matrix whose blocks tau2dgod (skkr_dbogo_bsf/clu/tau2d_god.f) builds
    by its recursion."""
    cfg = system.cfg
    lk, s, P = cfg.lk, cfg.s, cfg.P
    n = lk * s
    tint, tl, tr = system.tinv(E)
    gi = fourier(system.R, system.G_intra, k)
    go = fourier(system.R, system.G_inter, k)
    m00, m01, m10 = pl_blocks(gi, go)
    it, em = int(itermax or cfg.itermax), float(errmax or cfg.errmax)
    y, _, st1 = dugo(m00, m01, m10, tl, "L", it, em)
    x, _, st2 = dugo(m00, m01, m10, tr, "R", it, em)
    if st1 or st2:
        raise RuntimeError("decimation failed")
    A = np.zeros((P * n, P * n), np.complex128)
    for p in range(P):
        sl = slice(p * n, (p + 1) * n)
        A[sl, sl] = m00
        for i in range(s):
            si = slice(p * n + i * lk, p * n + (i + 1) * lk)
            A[si, si] += tint[p * s + i]
        if p + 1 < P:
            A[sl, (p + 1) * n:(p + 2) * n] = m01
            A[(p + 1) * n:(p + 2) * n, sl] = m10
    A[:n, :n] -= y
    A[(P - 1) * n:, (P - 1) * n:] -= x
    T = np.linalg.inv(A)
    ns = P * s
    return T.reshape(ns, lk, ns, lk).transpose(0, 2, 1, 3).copy()


def tau_pairs(system, E, ks, ws, pairs, a1, a2):
    """BZ sums of the off-diagonal tau of interface-site pairs as the cluster
    code accumulates them (skkr_dbogo_bsf/clu/cpacoord_save.f:432-478 with the
    identity point operation): pair (li, lj, n1, n2), 0-based sites,
    R = n1 a1 + n2 a2:  tau_ij = sum_k w e^{+i k.R} tau_{li,lj}(k),
    tau_ji = sum_k w e^{-i k.R} tau_{lj,li}(k)."""
    lk = system.cfg.lk
    tij = np.zeros((len(pairs), lk, lk), np.complex128)
    tji = np.zeros_like(tij)
    a1, a2 = np.asarray(a1, float), np.asarray(a2, float)
    for k, w in zip(ks, ws):
        T = full_tau_k(system, E, k)
        for q, (li, lj, n1, n2) in enumerate(pairs):
            R = n1 * a1 + n2 * a2
            ph = np.exp(1j * (k @ R))
            tij[q] += w * ph * T[li, lj]
            tji[q] += w * np.conj(ph) * T[lj, li]
    return tij, tji


def clebsch(lmax=5):
    """Clebsch-Gordan rectangular matrices of clebsch.f (lm -> kappa-mu):
    u1, u2 and 0-based index maps ind1, ind2 over kappa-mu indices
    ir = 2 kap^2 + kap + m (+1), l <= 5 (clebsch.f:21-67)."""
    u1, u2 = np.zeros(72), np.zeros(72)
    i1, i2 = np.zeros(72, np.int64), np.zeros(72, np.int64)
    inr = 0
    for l in range(6):
        t = float(2 * l + 1)
        for m in range(-l, l + 1):
            inr += 1
            if l != 0:
                kap = l
                ir = 2 * kap * kap + kap + m
                u1[ir - 1], i1[ir - 1] = np.sqrt((l + m) / t), inr
                ir += 1
                u2[ir - 1], i2[ir - 1] = -np.sqrt((l - m) / t), inr
            kap = -l - 1
            ir = 2 * kap * kap + kap + m
            u1[ir - 1], i1[ir - 1] = np.sqrt((l - m + 1) / t), inr
            ir += 1
            u2[ir - 1], i2[ir - 1] = np.sqrt((l + m + 1) / t), inr
    i1[i1 == 0] = 1
    i2[i2 == 0] = 1
    return u1, u2, i1 - 1, i2 - 1


def relmtrx(a, fac, lmax):
    """relmtrx.f:2-45: b(i, j) = u1(i) (fac a)(ind1 i, ind1 j) u1(j)
    + u2(i) (fac a)(ind2 i, ind2 j) u2(j), i, j over 2 (lmax+1)^2."""
    u1, u2, i1, i2 = clebsch()
    n, nl2 = 2 * (lmax + 1) ** 2, (lmax + 1) ** 2
    # zero-padded like relmtrx.f's a1(36,36): entries with u = 0 may point
    # past the lm block (clebsch.f reuses an index of the next l)
    a1 = np.zeros((36, 36), np.complex128)
    a1[:nl2, :nl2] = fac * np.asarray(a, np.complex128)
    b = np.zeros((n, n), np.complex128)
    for i in range(n):
        for j in range(n):
            b[i, j] = u1[i] * a1[i1[i], i1[j]] * u1[j] + u2[i] * a1[i2[i], i2[j]] * u2[j]
    return b
