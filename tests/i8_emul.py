"""CPU emulation of the INT8 digit complex product (zgemm_i8.h), step for
step, so that the GPU result can be compared bit for bit.

No reference counterpart: the reference multiplies in FP64 only
(skkr_kpar_v6/main/matops.f:204-228).  The emulation is pinned by the closed
forms in tests/test_i8_emul.py (digit representation, the magic-number digit
cut, error bound against the exact product, int32 bounds of the group sums,
exact products of small integers).

Per K-chunk of 32: each row of A and each column of B gets the exponent e
with max(|re|, |im|) < 2^e (0 for an all-zero line, at least -960); a value
becomes V = rint(x 2^(8S-2-e)) and its balanced base-256 digits
d_t = byte t of (V + sum_t 128 256^t) - 128.  The digit products with
(S-1-tA) + (S-1-tB) = g <= S-1 are summed per group g (exactly), two groups
share P_h = 256 G_2h + G_2h+1, and C += P_h 2^(eA + eB - 20 - 16h) for h
ascending, the real part as (P(re re) - P(im im)) and the imaginary part as
P(re im + im re).
"""
from __future__ import annotations

import numpy as np

KC = 32


def chunk_exponents(x: np.ndarray, axis: int) -> np.ndarray:
    """e with max(|re|, |im|) over `axis` < 2^e (0 for an all-zero line),
    at least -960."""
    mx = np.maximum(np.abs(x.real), np.abs(x.imag)).max(axis=axis)
    _, e = np.frexp(mx)
    e = np.where(mx > 0, e, 0).astype(np.int64)
    return np.maximum(e, -960)


def to_int(x: np.ndarray, e: np.ndarray, S: int) -> np.ndarray:
    """V = rint(x 2^(8S-2-e)) as int64 (e broadcast against x)."""
    return np.rint(np.ldexp(x, 8 * S - 2 - e)).astype(np.int64)


def digits(V: np.ndarray, S: int) -> np.ndarray:
    """Balanced base-256 digits of V, shape (S,) + V.shape, least significant
    first: V = sum_t d_t 256^t, d_t in [-128, 127]."""
    c = sum(128 << (8 * t) for t in range(S))
    W = V + c
    return np.stack([((W >> (8 * t)) & 255) - 128 for t in range(S)])


def magic_digits(x: np.ndarray, e: np.ndarray, S: int) -> np.ndarray:
    """The kernel's digit cut: D = x 2^(8S-2-e) + (1.5 2^52 + c) in FP64 (the
    product is exact, so one rounding as in the FMA), digit t = byte t of D's
    bit pattern XOR 0x80, read as a signed byte."""
    c = sum(128 << (8 * t) for t in range(S))
    M = np.float64(1.5 * 2.0 ** 52 + c)
    D = np.ldexp(x, 8 * S - 2 - e) + M
    bits = D.view(np.uint64)
    out = []
    for t in range(S):
        b = ((bits >> np.uint64(8 * t)) & np.uint64(255)).astype(np.int64) ^ 0x80
        out.append(np.where(b >= 128, b - 256, b))
    return np.stack(out)


def i8_product(A: np.ndarray, B: np.ndarray, S: int, return_groups=False):
    """C = A @ B as the INT8 digit kernel computes it (complex128)."""
    A = np.asarray(A, np.complex128)
    B = np.asarray(B, np.complex128)
    m, k = A.shape
    n = B.shape[1]
    cre = np.zeros((m, n))
    cim = np.zeros((m, n))
    pmax = 0
    H = (S + 1) // 2
    for k0 in range(0, k, KC):
        a = np.zeros((m, KC), np.complex128)
        b = np.zeros((KC, n), np.complex128)
        kk = min(KC, k - k0)
        a[:, :kk] = A[:, k0:k0 + kk]
        b[:kk, :] = B[k0:k0 + kk, :]
        ea = chunk_exponents(a, 1)
        eb = chunk_exponents(b, 0)
        ar = digits(to_int(a.real, ea[:, None], S), S).astype(np.float64)
        ai = digits(to_int(a.imag, ea[:, None], S), S).astype(np.float64)
        br = digits(to_int(b.real, eb[None, :], S), S).astype(np.float64)
        bi = digits(to_int(b.imag, eb[None, :], S), S).astype(np.float64)
        # group sums (integer products summed exactly in float64: |sum| < 2^31)
        gp = np.zeros((S + 1, m, n))
        gm = np.zeros((S + 1, m, n))
        gi = np.zeros((S + 1, m, n))
        for g in range(S):
            for i in range(g + 1):
                ta, tb = S - 1 - i, S - 1 - (g - i)
                gp[g] += ar[ta] @ br[tb]
                gm[g] += ai[ta] @ bi[tb]
                gi[g] += ar[ta] @ bi[tb] + ai[ta] @ br[tb]
        for h in range(H):
            pp = 256.0 * gp[2 * h] + gp[2 * h + 1]
            pm = 256.0 * gm[2 * h] + gm[2 * h + 1]
            pi = 256.0 * gi[2 * h] + gi[2 * h + 1]
            pmax = max(pmax, np.abs(pp).max(), np.abs(pm).max(), np.abs(pi).max())
            e = ea[:, None] + eb[None, :] - 20 - 16 * h
            cre = cre + np.ldexp(pp - pm, e)
            cim = cim + np.ldexp(pi, e)
    C = cre + 1j * cim
    if return_groups:
        return C, pmax
    return C


def error_bound(A: np.ndarray, B: np.ndarray, S: int) -> np.ndarray:
    """Elementwise bound of |i8_product - A @ B| (re and im separately).
    Per chunk and real product: rounding each operand value to V costs at
    most 2^(e - 8S + 1) per value, so |ab - VaVb 2^..| <= 2^(eA+eB) 2^(2-8S)
    per term; the dropped groups g = S .. 2S-2 hold at most 2S-1-g pairs of
    K products below 2^14 each, weighted 2^(eA+eB-12-8g): together below
    S K 2^(eA+eB) 2^(2-8S) 1.004.  Two real products per component."""
    m, k = A.shape
    n = B.shape[1]
    bound = np.zeros((m, n))
    for k0 in range(0, k, KC):
        a = A[:, k0:k0 + KC]
        b = B[k0:k0 + KC, :]
        ea = chunk_exponents(a, 1)
        eb = chunk_exponents(b, 0)
        kk = a.shape[1]
        rel = 2.0 ** (2 - 8 * S) * (1.0 + 1.004 * S)
        bound += 2 * kk * rel * np.ldexp(1.0, ea[:, None] + eb[None, :])
    return bound
