"""ctypes binding of libskkr_gpu.so (the C-ABI declared in include/skkr_gpu.h).

``Context`` mirrors the reference's k-loop contract (cpacoord's k-loop body,
skkr_kpar_v6/rsp-scf/cpacoord_ksave.f:188-400): set the structure constants
and t^-1, then ``tau_bz`` accumulates the BZ-integrated site-diagonal tau.
The per-routine methods (``gjinv``, ``surfgfd``, ``dugo``, ``tau2d``,
``pl_blocks``, ``zgemm``) stand for gjinv.f, surfgf.f, dugo.f, tau2d.f,
gstore+packer and doubmt2 and raise on the error codes where the reference
would stop.

There is no CPU fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .systems import colmajor, from_colmajor

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libskkr_gpu.so")

SKKR_OK = 0
ERRORS = {1: "SKKR_ERR_ARG", 2: "SKKR_ERR_HIP", 3: "SKKR_ERR_NOMEM",
          4: "SKKR_ERR_SINGULAR", 5: "SKKR_ERR_SGF_DIVERGED",
          6: "SKKR_ERR_SGF_ITERMAX", 7: "SKKR_ERR_UNSUPPORTED", 8: "SKKR_ERR_STR_NOCONV",
          9: "SKKR_ERR_TOO_LARGE"}
FP64, MIXED_FP32, MIXED_BF16, MIXED_INT8 = 0, 1, 2, 3

# exported symbols (every declaration of include/skkr_gpu.h)
SYMBOLS = ("skkr_create", "skkr_destroy", "skkr_error_string", "skkr_set_structure",
           "skkr_set_tinv", "skkr_tau_bz", "skkr_pl_blocks", "skkr_gjinv_batch",
           "skkr_zgemm_batch", "skkr_surfgfd_batch", "skkr_dugo_batch",
           "skkr_tau2d_batch", "skkr_set_option", "skkr_autotune", "skkr_set_screening",
           "skkr_tau_bz_phys", "skkr_str_create", "skkr_str_destroy", "skkr_gstore_batch",
           "skkr_set_structure_kkr", "skkr_cpa_bz", "skkr_set_pairs", "skkr_tau_bz_pairs",
           "skkr_set_structure_kkr_rel", "skkr_tau2d_layered_batch", "skkr_set_point_group",
           "skkr_surfgfd_noalloc", "skkr_tau2d_g2d", "skkr_dugo_g2d")


# codes that name a failing k-point or matrix (reported per item in status
# arrays, as the reference stops there); the others are call errors
PER_ITEM = (SKKR_OK, 4, 5, 6, 8)


class SkkrError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: {ERRORS.get(code, code)}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [("lk", ctypes.c_int), ("sites_per_pl", ctypes.c_int),
                ("n_pl", ctypes.c_int), ("errmax", ctypes.c_double),
                ("itermax_l", ctypes.c_int), ("itermax_r", ctypes.c_int),
                ("precision", ctypes.c_int), ("fp64_dec_below_imE", ctypes.c_double),
                ("device", ctypes.c_int), ("batch", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("ms_total", ctypes.c_double), ("ms_fourier", ctypes.c_double),
                ("ms_decimation", ctypes.c_double), ("ms_sweeps", ctypes.c_double),
                ("ms_accumulate", ctypes.c_double), ("ms_refine", ctypes.c_double),
                ("mean_iter_l", ctypes.c_double), ("mean_iter_r", ctypes.c_double),
                ("max_iter", ctypes.c_int), ("n_batches", ctypes.c_int),
                ("n_launches", ctypes.c_longlong), ("timing_detail", ctypes.c_int),
                ("dom_launches", ctypes.c_longlong), ("dom_ms", ctypes.c_double),
                ("dom_flops", ctypes.c_double), ("dom_name", ctypes.c_char * 32),
                ("dom_bytes", ctypes.c_double), ("dom_flops_exec", ctypes.c_double),
                ("inv_launches", ctypes.c_longlong), ("inv_ms", ctypes.c_double),
                ("inv_flops", ctypes.c_double), ("inv_bytes", ctypes.c_double),
                ("redo_batches", ctypes.c_int)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["dom_name"] = self.dom_name.decode()
        return d


class Tuning(ctypes.Structure):
    _fields_ = [("tile_nn", ctypes.c_int), ("tile_2nn", ctypes.c_int),
                ("us_nn", ctypes.c_double * 3), ("us_2nn", ctypes.c_double * 3)]

    def as_dict(self):
        return {"tile_nn": self.tile_nn, "tile_2nn": self.tile_2nn,
                "us_nn": list(self.us_nn), "us_2nn": list(self.us_2nn)}


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        V, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.skkr_create.argtypes = [ctypes.POINTER(V), ctypes.POINTER(Config)]
        L.skkr_destroy.argtypes = [V]
        L.skkr_destroy.restype = None
        L.skkr_error_string.argtypes = [I]
        L.skkr_error_string.restype = ctypes.c_char_p
        L.skkr_set_structure.argtypes = [V, I, V, V, V]
        L.skkr_set_tinv.argtypes = [V, V, V, V]
        L.skkr_tau_bz.argtypes = [V, D, D, I, V, V, V, V, V, V, ctypes.POINTER(Stats)]
        L.skkr_pl_blocks.argtypes = [V, I, V, V, V, V]
        L.skkr_gjinv_batch.argtypes = [V, I, I, V, V]
        L.skkr_zgemm_batch.argtypes = [V, I, I, V, V, V]
        L.skkr_surfgfd_batch.argtypes = [V, I, V, V, V, V, V, V]
        L.skkr_dugo_batch.argtypes = [V, ctypes.c_char, I, V, V, V, V, V, V]
        L.skkr_tau2d_batch.argtypes = [V, I, V, V, V, V, V, D, V, V]
        L.skkr_tau2d_layered_batch.argtypes = [V, I, V, V, V, V, V, D, V, V]
        L.skkr_set_point_group.argtypes = [V, I, V, V]
        L.skkr_set_option.argtypes = [V, ctypes.c_char_p, I]
        L.skkr_autotune.argtypes = [V, ctypes.POINTER(Tuning)]
        L.skkr_set_screening.argtypes = [V, I, V, V]
        L.skkr_tau_bz_phys.argtypes = [V, D, D, I, V, V, V, V, V, V, V, ctypes.POINTER(Stats)]
        L.skkr_str_create.argtypes = [ctypes.POINTER(V), ctypes.POINTER(StrConfig), V]
        L.skkr_str_destroy.argtypes = [V]
        L.skkr_str_destroy.restype = None
        L.skkr_gstore_batch.argtypes = [V, D, D, V, I, V, V, V]
        L.skkr_set_structure_kkr.argtypes = [V, V, V]
        L.skkr_set_structure_kkr_rel.argtypes = [V, V, V, I]
        L.skkr_cpa_bz.argtypes = [V, D, D, I, V, V, V, V, V, V, I, I, D, V, V, V, V, V, V, V,
                                  ctypes.POINTER(Stats)]
        L.skkr_set_pairs.argtypes = [V, I, V, V, V]
        L.skkr_tau_bz_pairs.argtypes = [V, D, D, I, V, V, V, V, V, V, V, ctypes.POINTER(Stats)]
        L.skkr_surfgfd_noalloc.argtypes = [I, I, I, V, V, V, V, V, V, D, I, V]
        L.skkr_tau2d_g2d.argtypes = [V, I, D, I, I, I, I, I, V, V, V, V, V, I, I, I, I, I, I, V, D, D]
        L.skkr_dugo_g2d.argtypes = [V, I, I, V, I, V, V, V, I, ctypes.c_char, I, I, V, I, I, I, V, D, D, V]
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c128(a):
    return np.ascontiguousarray(a, np.complex128)


class StrConfig(ctypes.Structure):
    _fields_ = [("lmax", ctypes.c_int), ("s", ctypes.c_int), ("nextra", ctypes.c_int),
                ("eta", ctypes.c_double), ("a1", ctypes.c_double * 2), ("a2", ctypes.c_double * 2),
                ("device", ctypes.c_int), ("batch", ctypes.c_int)]


class Context:
    """One device context (skkr_create ... skkr_destroy)."""

    def __init__(self, lk, sites_per_pl, n_pl, errmax=1e-8, itermax=50,
                 precision=FP64, fp64_dec_below_imE=0.02, device=0, batch=0):
        self.lk, self.s, self.P = lk, sites_per_pl, n_pl
        self.nl = 0  # trace groups of the screening (set_screening)
        self.n = lk * sites_per_pl
        cfg = Config(lk, sites_per_pl, n_pl, errmax, itermax, itermax, precision,
                     fp64_dec_below_imE, device, batch)
        h = ctypes.c_void_p()
        self._check(lib().skkr_create(ctypes.byref(h), ctypes.byref(cfg)), "skkr_create")
        self._h = h

    @classmethod
    def for_system(cls, system, **kw):
        c = system.cfg
        kw.setdefault("errmax", c.errmax)
        kw.setdefault("itermax", c.itermax)
        ctx = cls(c.lk, c.s, c.P, **kw)
        ctx.set_structure(system.R, system.G_intra, system.G_inter)
        return ctx

    @staticmethod
    def _check(rc, what):
        if rc != SKKR_OK:
            raise SkkrError(rc, what)

    @staticmethod
    def _check_call(rc, what):
        """Raise on call errors (argument, HIP, memory, unsupported); per-item
        failures stay in the returned status arrays."""
        if rc not in PER_ITEM:
            raise SkkrError(rc, what)

    def close(self):
        if getattr(self, "_h", None):
            lib().skkr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_structure_kkr(self, structure, alpha, rel=0):
        """Take the structure constants of the k-loop from a ``Structure``
        (the layer producer) with screening alpha[layer, l] at the energy of
        the next tau_bz calls; ``structure=None`` returns to the Fourier sum.
        rel 1 / 2: kappa-mu basis (lk = 2 (lmax+1)^2), constants at
        E + E^2/c^2 / at E."""
        if structure is None:
            self._check(lib().skkr_set_structure_kkr(self._h, None, None), "skkr_set_structure_kkr")
            self._str = None
            return
        al = _c128(alpha)
        if al.shape != (structure.s, structure.lmax + 1):
            raise ValueError("alpha must be s x (lmax+1)")
        self._check(lib().skkr_set_structure_kkr_rel(self._h, structure._h, _p(al), int(rel)),
                    "skkr_set_structure_kkr_rel")
        self._str = structure  # keeps the producer alive while bound

    def cpa_bz(self, E, k, w, conc, tinv_a, tinv_b, tinv_c=None, itcpam=50, cpatol=1e-10,
               raise_on_error=True):
        """CPA loop of one energy (skkr_cpa_bz): conc[site], tinv_a/b
        [site, lk, lk]; tinv_c None starts from the average t-matrix (tata).
        Returns tau_c, tau_a, tau_b, the coherent tinv_c, passes and error."""
        k = np.ascontiguousarray(k, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        nk = len(w)
        ns = self.P * self.s
        shp = (ns, self.lk, self.lk)
        cc = np.ascontiguousarray(conc, np.float64)
        a, b = colmajor(_c128(tinv_a)), colmajor(_c128(tinv_b))
        mc = colmajor(_c128(tinv_c) if tinv_c is not None else np.zeros(shp, np.complex128))
        tc, ta, tb = (colmajor(np.zeros(shp, np.complex128)) for _ in range(3))
        itc = ctypes.c_int(0)
        err = ctypes.c_double(0.0)
        iters = np.zeros((nk, 2), np.int32)
        status = np.zeros(nk, np.int32)
        st = Stats()
        rc = lib().skkr_cpa_bz(self._h, float(np.real(E)), float(np.imag(E)), nk, _p(k), _p(w), _p(cc),
                               _p(a), _p(b), _p(mc), 1 if tinv_c is None else 0, int(itcpam),
                               float(cpatol), _p(tc), _p(ta), _p(tb), ctypes.byref(itc),
                               ctypes.byref(err), _p(iters), _p(status), ctypes.byref(st))
        if raise_on_error:
            self._check(rc, "skkr_cpa_bz")
        return {"tau_c": from_colmajor(tc), "tau_a": from_colmajor(ta), "tau_b": from_colmajor(tb),
                "tinv_c": from_colmajor(mc), "itcpa": itc.value, "error": err.value, "iters": iters,
                "status": status, "rc": rc, "stats": st.as_dict()}

    def set_pairs(self, pairs, a1, a2):
        """Site pairs (li, lj, n1, n2), 0-based sites, of the off-diagonal sums."""
        pr = np.array([(li + 1, lj + 1, n1, n2) for li, lj, n1, n2 in pairs], np.int32).reshape(-1)
        A1, A2 = np.ascontiguousarray(a1, np.float64), np.ascontiguousarray(a2, np.float64)
        self._check(lib().skkr_set_pairs(self._h, len(pairs), _p(pr), _p(A1), _p(A2)), "skkr_set_pairs")
        self._npair = len(pairs)

    def tau_bz_pairs(self, E, k, w):
        """skkr_tau_bz plus the pair sums: returns tau_bz, tau_ij, tau_ji."""
        k = np.ascontiguousarray(k, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        nk = len(w)
        ns = self.P * self.s
        acc = colmajor(np.zeros((ns, self.lk, self.lk), np.complex128))
        tij = colmajor(np.zeros((self._npair, self.lk, self.lk), np.complex128))
        tji = colmajor(np.zeros((self._npair, self.lk, self.lk), np.complex128))
        iters = np.zeros((nk, 2), np.int32)
        status = np.zeros(nk, np.int32)
        st = Stats()
        rc = lib().skkr_tau_bz_pairs(self._h, float(np.real(E)), float(np.imag(E)), nk, _p(k), _p(w),
                                     _p(acc), _p(tij), _p(tji), _p(iters), _p(status), ctypes.byref(st))
        self._check(rc, "skkr_tau_bz_pairs")
        return {"tau_bz": from_colmajor(acc), "tau_ij": from_colmajor(tij), "tau_ji": from_colmajor(tji),
                "iters": iters, "status": status, "stats": st.as_dict()}

    def set_option(self, key, value):
        self._check(lib().skkr_set_option(self._h, key.encode(), int(value)),
                    f"skkr_set_option({key})")

    def autotune(self):
        t = Tuning()
        self._check(lib().skkr_autotune(self._h, ctypes.byref(t)), "skkr_autotune")
        return t.as_dict()

    # ---- k-loop -----------------------------------------------------------
    def set_structure(self, R, G_intra, G_inter):
        R = np.ascontiguousarray(R, np.float64)
        self._gi = colmajor(_c128(G_intra))
        self._go = colmajor(_c128(G_inter))
        self._check(lib().skkr_set_structure(self._h, len(R), _p(R), _p(self._gi), _p(self._go)),
                    "skkr_set_structure")

    def set_tinv(self, tinv_int, tinv_l, tinv_r):
        a, b, c = (colmajor(_c128(t)) for t in (tinv_int, tinv_l, tinv_r))
        self._check(lib().skkr_set_tinv(self._h, _p(a), _p(b), _p(c)), "skkr_set_tinv")

    def tau_bz(self, E, k, w, tau_bz=None, want_trk=True, timing_detail=False,
               raise_on_error=True):
        k = np.ascontiguousarray(k, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        nk = len(w)
        ns = self.P * self.s
        acc = np.zeros((ns, self.lk, self.lk), np.complex128) if tau_bz is None else tau_bz
        buf = colmajor(acc)
        trk = np.zeros((nk, ns), np.complex128) if want_trk else None
        iters = np.zeros((nk, 2), np.int32)
        status = np.zeros(nk, np.int32)
        st = Stats()
        st.timing_detail = 1 if timing_detail else 0
        rc = lib().skkr_tau_bz(self._h, float(np.real(E)), float(np.imag(E)), nk, _p(k), _p(w),
                               _p(buf), _p(trk), _p(iters), _p(status), ctypes.byref(st))
        if raise_on_error:
            self._check(rc, "skkr_tau_bz")
        acc[...] = from_colmajor(buf)
        return {"tau_bz": acc, "trk": trk, "iters": iters, "status": status,
                "stats": st.as_dict(), "rc": rc}

    def set_screening(self, a, l_of_index, nl=None):
        """a: (sites, lk) complex screening per basis index; l_of_index: (lk,)."""
        a = np.ascontiguousarray(a, np.complex128)
        li = np.ascontiguousarray(l_of_index, np.int32)
        nl = int(nl if nl is not None else li.max() + 1)
        self._check(lib().skkr_set_screening(self._h, nl, _p(li), _p(a)), "skkr_set_screening")
        self.nl = nl

    def tau_bz_phys(self, E, k, w, want_taukdiag=True, raise_on_error=True):
        """skkr_tau_bz with the physical representation: returns tau_bz,
        tau_bz_phys (sites, lk, lk) and taukdiag (nk, sites, nl)."""
        k = np.ascontiguousarray(k, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        nk = len(w)
        ns = self.P * self.s
        acc = np.zeros((ns, self.lk, self.lk), np.complex128)
        buf, pbuf = colmajor(acc), colmajor(acc)
        tkd = np.zeros((ns, self.nl, nk), np.complex128) if want_taukdiag else None
        iters = np.zeros((nk, 2), np.int32)
        status = np.zeros(nk, np.int32)
        st = Stats()
        rc = lib().skkr_tau_bz_phys(self._h, float(np.real(E)), float(np.imag(E)), nk, _p(k),
                                    _p(w), _p(buf), _p(pbuf), _p(tkd), _p(iters), _p(status),
                                    ctypes.byref(st))
        if raise_on_error:
            self._check(rc, "skkr_tau_bz_phys")
        out = {"tau_bz": from_colmajor(buf), "tau_bz_phys": from_colmajor(pbuf), "iters": iters,
               "status": status, "stats": st.as_dict(), "rc": rc}
        if tkd is not None:
            out["taukdiag"] = np.ascontiguousarray(np.transpose(tkd, (2, 0, 1)))
        return out

    # ---- per-routine entry points ----------------------------------------
    def pl_blocks(self, k):
        k = np.ascontiguousarray(k, np.float64)
        nk = len(k)
        out = [np.zeros((nk, self.n, self.n), np.complex128) for _ in range(3)]
        self._check(lib().skkr_pl_blocks(self._h, nk, _p(k), *map(_p, out)), "skkr_pl_blocks")
        return tuple(from_colmajor(o) for o in out)

    def gjinv(self, a):
        a = np.asarray(a)
        nb, n = a.shape[0], a.shape[1]
        buf = colmajor(_c128(a))
        status = np.zeros(nb, np.int32)
        rc = lib().skkr_gjinv_batch(self._h, n, nb, _p(buf), _p(status))
        self._check_call(rc, "skkr_gjinv_batch")
        return from_colmajor(buf), status

    def zgemm(self, a, b):
        nb, n = a.shape[0], a.shape[1]
        A, B = colmajor(_c128(a)), colmajor(_c128(b))
        C = np.zeros_like(A)
        self._check(lib().skkr_zgemm_batch(self._h, n, nb, _p(A), _p(B), _p(C)), "skkr_zgemm_batch")
        return from_colmajor(C)

    def surfgfd(self, ml, m0, mr):
        nb = ml.shape[0]
        bufs = [colmajor(_c128(m)) for m in (ml, m0, mr)]
        x = np.zeros_like(bufs[0])
        iters = np.zeros(nb, np.int32)
        status = np.zeros(nb, np.int32)
        self._check_call(lib().skkr_surfgfd_batch(self._h, nb, *map(_p, bufs), _p(x), _p(iters), _p(status)),
                         "skkr_surfgfd_batch")
        return from_colmajor(x), iters, status

    def dugo(self, side, m00, m01, m10):
        nb = m00.shape[0]
        bufs = [colmajor(_c128(m)) for m in (m00, m01, m10)]
        xy = np.zeros_like(bufs[0])
        iters = np.zeros(nb, np.int32)
        status = np.zeros(nb, np.int32)
        self._check_call(lib().skkr_dugo_batch(self._h, side.encode(), nb, *map(_p, bufs), _p(xy),
                                               _p(iters), _p(status)), "skkr_dugo_batch")
        return from_colmajor(xy), iters, status

    def tau2d(self, m00, m01, m10, x, y, w=1.0):
        nb = m00.shape[0]
        bufs = [colmajor(_c128(m)) for m in (m00, m01, m10, x, y)]
        ns = self.P * self.s
        tau = np.zeros((nb, ns, self.lk, self.lk), np.complex128)
        tb = colmajor(tau)
        status = np.zeros(nb, np.int32)
        self._check_call(lib().skkr_tau2d_batch(self._h, nb, *map(_p, bufs), float(w), _p(tb), _p(status)),
                         "skkr_tau2d_batch")
        return from_colmajor(tb), status

    def set_point_group(self, dg=None, dg1=None):
        """Point group of an irreducible-wedge mesh (cpacoord.f:294-299):
        dg, dg1 (ng, lk, lk); None switches it off."""
        if dg is None:
            self._check(lib().skkr_set_point_group(self._h, 0, None, None), "skkr_set_point_group")
            return
        d, d1 = colmajor(_c128(dg)), colmajor(_c128(dg1))
        if d.shape != d1.shape or d.shape[1:] != (self.lk, self.lk):
            raise ValueError("dg, dg1 must be (ng, lk, lk)")
        self._check(lib().skkr_set_point_group(self._h, len(d), _p(d), _p(d1)), "skkr_set_point_group")

    def tau2d_layered(self, m00, m01, m10, x, y, w=1.0):
        """tau2d of a layered stack (bulkgeo false): m00, m01, m10 are
        (nb, P, n, n), the packer blocks of every PL's own g2d plane; x, y
        (nb, n, n) the lead terms.  Returns (tau (nb, P*s, lk, lk), status);
        raises on an argument error."""
        nb = m00.shape[0]
        if m00.shape[:2] != (nb, self.P) or m01.shape != m00.shape or m10.shape != m00.shape:
            raise ValueError("m00, m01, m10 must be (nb, P, n, n)")
        bufs = [colmajor(_c128(m)) for m in (m00, m01, m10, x, y)]
        tau = np.zeros((nb, self.P * self.s, self.lk, self.lk), np.complex128)
        tb = colmajor(tau)
        status = np.zeros(nb, np.int32)
        rc = lib().skkr_tau2d_layered_batch(self._h, nb, *map(_p, bufs), float(w), _p(tb), _p(status))
        self._check_call(rc, "skkr_tau2d_layered_batch")
        return from_colmajor(tb), status

    # -- the reference's own argument lists (one k-point per call) ----------
    @staticmethod
    def _fortran(a, shape):
        a = np.asfortranarray(a, np.complex128)
        if a.shape != tuple(shape):
            raise ValueError(f"array of shape {a.shape}, expected {tuple(shape)}")
        return a

    def tau2d_g2d(self, irel, wgeff, bulkgeo, lmax, tminv, g2d, x, y, tau, ninprc, fac=1.0):
        """tau2d(irel, wgeff, bulkgeo, lmax, lkmax, lkmaxp, nintfc, tminv, g2d,
        x, y, tau, mdimp) of skkr_kpar_v6/main/tau2d.f:2-3 on Fortran-ordered
        arrays: tminv and tau (lkmaxp, lkmaxp, mintfc), g2d (lmmaxp, lmmaxp,
        2 minprc + 1, minprc, mprc1 + 2), x and y (mdimp, mdimp); ninprc is
        common /lay2d/'s ninprc(0:nprc+1).  tau is written in place (its first
        nintfc layers).  Returns the status code."""
        lkmaxp, _, mintfc = tminv.shape
        lmmaxp, _, nj, minprc, npl = g2d.shape
        mdimp = x.shape[0]
        ninprc = np.ascontiguousarray(ninprc, np.int32)
        nprc = len(ninprc) - 2
        nintfc = int(ninprc[1:nprc + 1].sum())
        tm = self._fortran(tminv, (lkmaxp, lkmaxp, mintfc))
        g = self._fortran(g2d, (lmmaxp, lmmaxp, 2 * minprc + 1, minprc, npl))
        xx, yy = self._fortran(x, (mdimp, mdimp)), self._fortran(y, (mdimp, mdimp))
        if not (isinstance(tau, np.ndarray) and tau.dtype == np.complex128 and tau.flags.f_contiguous
                and tau.shape == (lkmaxp, lkmaxp, mintfc)):
            raise ValueError("tau must be a Fortran-ordered complex128 (lkmaxp, lkmaxp, mintfc) array")
        rc = lib().skkr_tau2d_g2d(self._h, irel, float(wgeff), int(bool(bulkgeo)), lmax, self.lk, lkmaxp, nintfc,
                                  _p(tm), _p(g), _p(xx), _p(yy), _p(tau), mdimp, lmmaxp, minprc, npl - 2,
                                  mintfc, nprc, _p(ninprc), complex(fac).real, complex(fac).imag)
        self._check_call(rc, "skkr_tau2d_g2d")
        return rc

    def dugo_g2d(self, lmax, tminvb, g2d, g2di, xy, side, bulk, irel, ninprc, fac=1.0):
        """dugo(lmax, lkmax, tminvb, lkmaxp, g2d, g2di, xy, mdimp, side, bulk,
        irel, sgfconv) of skkr_kpar_v6/main/dugo.f:9-10 on Fortran-ordered
        arrays: tminvb (lkmaxp, lkmaxp, minprc), g2d and g2di (lmmaxp, lmmaxp,
        2 minprc + 1, minprc), xy (mdimp, mdimp), written in place.  Returns
        (status code, sgfconv, iterations)."""
        lkmaxp = tminvb.shape[0]
        lmmaxp, _, nj, minprc = g2d.shape
        mdimp = xy.shape[0]
        ninprc = np.ascontiguousarray(ninprc, np.int32)
        nprc = len(ninprc) - 2
        tb = self._fortran(tminvb, (lkmaxp, lkmaxp, tminvb.shape[2]))
        g = self._fortran(g2d, (lmmaxp, lmmaxp, 2 * minprc + 1, minprc))
        gi = self._fortran(g2di, (lmmaxp, lmmaxp, 2 * minprc + 1, minprc))
        if not (isinstance(xy, np.ndarray) and xy.dtype == np.complex128 and xy.flags.f_contiguous
                and xy.shape == (mdimp, mdimp)):
            raise ValueError("xy must be a Fortran-ordered complex128 (mdimp, mdimp) array")
        conv = ctypes.c_int(0)
        its = ctypes.c_int(0)
        rc = lib().skkr_dugo_g2d(self._h, lmax, self.lk, _p(tb), lkmaxp, _p(g), _p(gi), _p(xy), mdimp,
                                 side.encode() if isinstance(side, str) else side, int(bool(bulk)), irel,
                                 ctypes.byref(conv), lmmaxp, minprc, nprc, _p(ninprc),
                                 complex(fac).real, complex(fac).imag, ctypes.byref(its))
        self._check_call(rc, "skkr_dugo_g2d")
        return rc, bool(conv.value), its.value


class Structure:
    """Layer structure constants on the GPU (skkr_str_create ...
    skkr_str_destroy): the reference's gstore (skkr_kpar_v6/main/gstore.f:6)
    for the bulk-like geometry, per k-point."""

    def __init__(self, lmax, s, nextra, eta, a1, a2, cvec, device=0, batch=0):
        self.lmax, self.s, self.nextra = lmax, s, nextra
        self.nl2 = (lmax + 1) ** 2
        cfg = StrConfig(lmax, s, nextra, eta, (ctypes.c_double * 2)(*a1), (ctypes.c_double * 2)(*a2),
                        device, batch)
        cv = np.ascontiguousarray(cvec, np.float64)
        if cv.shape != (s + 1, 3):
            raise ValueError("cvec must be (s+1) x 3")
        h = ctypes.c_void_p()
        Context._check(lib().skkr_str_create(ctypes.byref(h), ctypes.byref(cfg), _p(cv)), "skkr_str_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().skkr_str_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def gstore(self, E, alpha, k, check=True):
        """g2d per k as an array [k, icpl, icount, jj+s, n, m] (the reverse of
        the reference's g2d(m, n, jj, icount, icpl)); alpha[layer, l]."""
        al = _c128(alpha)
        if al.shape != (self.s, self.lmax + 1):
            raise ValueError("alpha must be s x (lmax+1)")
        k = np.ascontiguousarray(k, np.float64).reshape(-1, 2)
        nk = len(k)
        s, nl2 = self.s, self.nl2
        out = np.zeros((nk, 3, s, 2 * s + 1, nl2, nl2), np.complex128)
        st = np.zeros(nk, np.int32)
        rc = lib().skkr_gstore_batch(self._h, float(np.real(E)), float(np.imag(E)), _p(al), nk,
                                     _p(k), _p(out), _p(st))
        if check:
            Context._check(rc, "skkr_gstore_batch")
        return out, st
