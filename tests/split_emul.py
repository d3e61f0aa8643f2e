"""CPU emulation of the split-BF16 complex product (zgemm_split.h), step for
step, so that the GPU result can be compared bit for bit.

No reference counterpart: the reference multiplies in FP64 only
(skkr_kpar_v6/main/matops.f:204-228).  The emulation is pinned by the closed
forms in tests/test_split_emul.py (error bound against the FP64 product,
exactness of the integer group sums, slice recombination).

Per K-chunk of 32: each row of A and each column of B gets the exponent e
with max(|re|, |im|) < 2^e; y = x 2^-e is cut into S slices
s_1 = rint(256 y), s_{i+1} = rint(512 r_i); the products of slices i and j
(0-based, i + j <= S - 1) are summed exactly per group g = i + j; the groups
are recombined as v = acc_0 + 2^-9 (acc_1 + 2^-9 (...)) and C += v 2^(eA+eB-16).
"""
from __future__ import annotations

import numpy as np

KC = 32


def chunk_exponents(x: np.ndarray, axis: int) -> np.ndarray:
    """e with max(|re|, |im|) over `axis` < 2^e (0 for an all-zero line)."""
    mx = np.maximum(np.abs(x.real), np.abs(x.imag)).max(axis=axis)
    _, e = np.frexp(mx)
    return np.where(mx > 0, e, 0).astype(np.int64)


def slices(y: np.ndarray, S: int) -> np.ndarray:
    """S integer slices of real y in (-1, 1): shape (S,) + y.shape."""
    out = np.empty((S,) + y.shape)
    y = y * 256.0
    for i in range(S):
        s = np.rint(y)
        out[i] = s
        y = (y - s) * 512.0
    return out


def split_product(A: np.ndarray, B: np.ndarray, S: int, return_groups=False):
    """C = A @ B as the split-BF16 kernel computes it (complex128)."""
    A = np.asarray(A, np.complex128)
    B = np.asarray(B, np.complex128)
    m, k = A.shape
    n = B.shape[1]
    cre = np.zeros((m, n))
    cim = np.zeros((m, n))
    gmax = 0.0
    for k0 in range(0, k, KC):
        a = np.zeros((m, KC), np.complex128)
        b = np.zeros((KC, n), np.complex128)
        kk = min(KC, k - k0)
        a[:, :kk] = A[:, k0:k0 + kk]
        b[:kk, :] = B[k0:k0 + kk, :]
        ea = chunk_exponents(a, 1)
        eb = chunk_exponents(b, 0)
        ya = np.ldexp(a.real, -ea[:, None]), np.ldexp(a.imag, -ea[:, None])
        yb = np.ldexp(b.real, -eb[None, :]), np.ldexp(b.imag, -eb[None, :])
        ar, ai = slices(ya[0], S), slices(ya[1], S)
        br, bi = slices(yb[0], S), slices(yb[1], S)
        gre = np.zeros((S, m, n))
        gim = np.zeros((S, m, n))
        for i in range(S):
            for j in range(S - i):
                # integer products: exact in float64
                gre[i + j] += ar[i] @ br[j] - ai[i] @ bi[j]
                gim[i + j] += ar[i] @ bi[j] + ai[i] @ br[j]
        gmax = max(gmax, np.abs(gre).max(), np.abs(gim).max())
        vr = gre[S - 1].copy()
        vi = gim[S - 1].copy()
        for g in range(S - 2, -1, -1):
            vr = vr * (1.0 / 512.0) + gre[g]
            vi = vi * (1.0 / 512.0) + gim[g]
        e = ea[:, None] + eb[None, :] - 16
        cre = cre + np.ldexp(vr, e)
        cim = cim + np.ldexp(vi, e)
    C = cre + 1j * cim
    if return_groups:
        return C, gmax
    return C


def error_bound(A: np.ndarray, B: np.ndarray, S: int) -> np.ndarray:
    """Elementwise bound of |split_product - A @ B| (re and im separately):
    per chunk, slicing error 2^-(9S) per operand value and the dropped slice
    pairs, times K_chunk * 2 * 2^(eA + eB)."""
    m, k = A.shape
    n = B.shape[1]
    bound = np.zeros((m, n))
    for k0 in range(0, k, KC):
        a = A[:, k0:k0 + KC]
        b = B[k0:k0 + KC, :]
        ea = chunk_exponents(a, 1)
        eb = chunk_exponents(b, 0)
        kk = a.shape[1]
        # |x - slices| <= 2^-9S relative to 2^e for each factor; dropped pairs
        # i + j >= S contribute at most S * 2^-(9S - 1) (|s| <= 256)
        rel = 2.0 * 2.0 ** (-9 * S) + S * 2.0 ** (-9 * S + 1)
        bound += 2 * kk * rel * np.ldexp(1.0, ea[:, None] + eb[None, :])
    return bound
