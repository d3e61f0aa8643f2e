"""Adaptive refinement of the 2D Brillouin-zone mesh for the BZ integral of tr G.

The integrand is f(k) = tr G(E, k) = sum over interface sites of tr tau(k),
the per-k trace the k-loop returns (``trk``); the integral is the BZ average,
the equal-weight normalisation of the reference's k-sets (kmesh2d.f:245-251).

Mesh.  The BZ [-pi, pi)^2 is triangulated from an N0 x N0 grid of points
(each square cell split into two triangles).  Subdividing a triangle into
four at its edge midpoints, applied to both triangles of a cell, splits the
cell into four cells of the next level, so the vertices of level l are the
points of the N0 2^l grid.  All levels are sub-lattices of the finest one
(M = N0 2^Lmax points per 2 pi): shared vertices and periodic images are
identified exactly and every k-point is evaluated once.

Quadrature on a partly refined mesh.  The vertex rule on a uniform level is
the periodic trapezoid rule, whose error falls exponentially with the level
for the analytic, periodic f -- much faster than the O(h^2) local error of a
triangle suggests, because the local errors cancel over the periodic domain.
A refined patch with a sharp boundary breaks that cancellation, so the
refined region enters through a smooth window instead (a partition of unity):

    I = sum_l Q_l[ f * phi_l * (1 - W_l) ] ,   phi_0 = 1,
    W_l = sum_{c in U_l} B_c ,   phi_{l+1} = phi_l * W_l ,

where Q_l is the trapezoid rule of level l, U_l the set of level-l cells
proposed for subdivision and B_c a smooth bump per cell (the cell's indicator
convolved with a Gaussian of width sigma cells, cut at cut*sigma), with
sum_c B_c = 1 over all cells.  The identity is exact for any proposal; the
proposal decides the accuracy.  Each term is a windowed, periodic function
that its level's grid resolves when U_l holds the cells whose bumps f * B_c
are not resolved at level l, measured by the indicator

    e_l(c) = | Q_{l+1}[f phi_l B_c] - Q_l[f phi_l B_c] |

(the aliasing of the bump, spectrally small where f is smooth).  Points of
level l are evaluated only where phi_l > 0.

Proposals:
  * ``uniform``: every cell (plain uniform refinement: the judge and the
    fallback);
  * ``probe``: the exact indicator, which needs level l+1 on the whole active
    region (the method without a network: it pays for the probe points);
  * ``learned``: an MLP predicts log10 e_l(c) / |I| from level-l data only
    (``cell_features``) and no probe points are evaluated.
"""
from __future__ import annotations

import math
import warnings

import numpy as np
from scipy.special import erf

FEATURES = ("log_abs_f", "log_grad", "log_d2", "log_d4", "log_d2_nb3", "log_d2_nb6",
            "log_imE", "log_h", "log_phi")
NFEAT = len(FEATURES)
LOG_FLOOR = -16.0
MAX_FINEST = 4096          # dense finest lattice (complex128) per side


def _bump_matrix(npts: int, ncell: int, sigma: float, cut: float) -> np.ndarray:
    """b[i, c]: periodic bump of cell c (ncell cells per 2 pi) at point i
    (npts per 2 pi), in cell units: the cell's indicator convolved with a
    Gaussian of width sigma, cut at cut*sigma, summed over periodic images.
    Rows sum to 1."""
    x = np.arange(npts) * (ncell / npts)
    c = np.arange(ncell)
    d = (x[:, None] - c[None, :] + ncell / 2) % ncell - ncell / 2
    lim = cut * sigma

    def T(u):
        v = 0.5 * (1.0 + erf(u / (math.sqrt(2.0) * sigma)))
        return np.where(u <= -lim, 0.0, np.where(u >= lim, 1.0, v))

    K = int(math.ceil((lim + 1.0) / ncell)) + 1
    b = np.zeros((npts, ncell))
    for m in range(-K, K + 1):
        b += T(d + m * ncell) - T(d + m * ncell - 1.0)
    return b


def _log(x):
    return np.log10(np.maximum(np.nan_to_num(x, nan=0.0), 10.0 ** LOG_FLOOR))


def _max_filter(a: np.ndarray, r: int) -> np.ndarray:
    """Periodic (2r+1)^2 maximum filter of a non-negative array."""
    out = a.copy()
    for ax in (0, 1):
        b = out.copy()
        for s in range(1, r + 1):
            b = np.maximum(b, np.maximum(np.roll(out, s, ax), np.roll(out, -s, ax)))
        out = b
    return out


class WindowedMesh:
    """Nested-lattice BZ mesh with the windowed multilevel trapezoid rule."""

    def __init__(self, evaluate, n0: int = 32, lmax: int = 6, im_e: float = 0.01,
                 sigma: float = 1.0, cut: float = 6.0):
        """evaluate(k: (m, 2) float array) -> (m,) complex f values."""
        self.evaluate = evaluate
        self.n0, self.lmax, self.im_e = n0, lmax, float(im_e)
        self.M = n0 << lmax
        if self.M > MAX_FINEST:
            raise ValueError(f"finest lattice {self.M} exceeds {MAX_FINEST} points per side")
        self.sigma, self.cut = sigma, cut
        self.grid = np.full((self.M, self.M), np.nan + 1j * np.nan, dtype=np.complex128)
        self.n_evals = 0
        self.flags = []            # U_l, dense bool (n_l, n_l)
        self._bm = {}

    # -- lattice bookkeeping -------------------------------------------
    def n_at(self, level: int) -> int:
        return self.n0 << level

    def _stride(self, level: int) -> int:
        return self.M // self.n_at(level)

    def bump(self, npts: int, ncell: int) -> np.ndarray:
        key = (npts, ncell)
        if key not in self._bm:
            self._bm[key] = _bump_matrix(npts, ncell, self.sigma, self.cut)
        return self._bm[key]

    def ensure(self, level: int, mask: np.ndarray) -> None:
        """Evaluate f at the level-`level` points selected by mask (once each)."""
        s = self._stride(level)
        sub = self.grid[::s, ::s]
        need = mask & np.isnan(sub.real)
        if not need.any():
            return
        ix, iy = np.nonzero(need)
        k = -math.pi + 2.0 * math.pi * np.stack([ix, iy], 1).astype(np.float64) / self.n_at(level)
        vals = np.asarray(self.evaluate(k), dtype=np.complex128)
        self.grid[ix * s, iy * s] = vals
        self.n_evals += len(ix)

    def values(self, level: int) -> np.ndarray:
        s = self._stride(level)
        return self.grid[::s, ::s]

    def window(self, upto: int, level: int) -> np.ndarray:
        """phi_upto = prod_{j < upto} W_j on the level-`level` grid."""
        n = self.n_at(level)
        phi = np.ones((n, n))
        for j in range(upto):
            B = self.bump(n, self.n_at(j))
            phi *= B @ self.flags[j].astype(np.float64) @ B.T
        phi[phi < 1e-15] = 0.0
        return phi

    # -- indicator and features ------------------------------------------
    def indicator(self, level: int, phi_l: np.ndarray) -> np.ndarray:
        """e_l(c) for all level-l cells (needs level l+1 where phi_l > 0)."""
        n = self.n_at(level)
        phi_f = self.window(level, level + 1)
        self.ensure(level + 1, phi_f > 0)
        g1 = np.where(phi_l > 0, self.values(level), 0.0) * phi_l
        g2 = np.where(phi_f > 0, self.values(level + 1), 0.0) * phi_f
        B1, B2 = self.bump(n, n), self.bump(2 * n, n)
        return np.abs(B2.T @ g2 @ B2 / (2 * n) ** 2 - B1.T @ g1 @ B1 / n ** 2)

    def cell_features(self, level: int, phi_l: np.ndarray, scale: float) -> np.ndarray:
        """Per-cell features from level-l values only: (n_l^2, NFEAT) float32.
        scale: |I| estimate the magnitudes are measured against."""
        n = self.n_at(level)
        f = np.where(phi_l > 0, self.values(level), np.nan)
        sh = lambda a, dx, dy: np.roll(np.roll(a, -dx, 0), -dy, 1)
        # corners of cell (i, j): points (i, j), (i+1, j), (i, j+1), (i+1, j+1)
        corners = np.stack([f, sh(f, 1, 0), sh(f, 0, 1), sh(f, 1, 1)])
        gx, gy = np.abs(sh(f, 1, 0) - f), np.abs(sh(f, 0, 1) - f)
        grad = np.fmax(np.fmax(gx, sh(gx, 0, 1)), np.fmax(gy, sh(gy, 1, 0)))
        d2x = np.abs(sh(f, 1, 0) - 2 * f + sh(f, -1, 0))
        d2y = np.abs(sh(f, 0, 1) - 2 * f + sh(f, 0, -1))
        d2p = np.nan_to_num(np.fmax(d2x, d2y), nan=0.0)
        d2 = np.maximum.reduce([d2p, sh(d2p, 1, 0), sh(d2p, 0, 1), sh(d2p, 1, 1)])
        d4x = np.abs(sh(f, 2, 0) - 4 * sh(f, 1, 0) + 6 * f - 4 * sh(f, -1, 0) + sh(f, -2, 0))
        d4y = np.abs(sh(f, 0, 2) - 4 * sh(f, 0, 1) + 6 * f - 4 * sh(f, 0, -1) + sh(f, 0, -2))
        d4p = np.nan_to_num(np.fmax(d4x, d4y), nan=0.0)
        d4 = np.maximum.reduce([d4p, sh(d4p, 1, 0), sh(d4p, 0, 1), sh(d4p, 1, 1)])
        phic = np.stack([phi_l, sh(phi_l, 1, 0), sh(phi_l, 0, 1), sh(phi_l, 1, 1)]).mean(0)
        with np.errstate(invalid="ignore"), warnings.catch_warnings():
            warnings.simplefilter("ignore", RuntimeWarning)
            absf = np.nanmean(np.abs(corners), axis=0)
        s = max(scale, 1e-300)
        feats = [_log(absf / s), _log(grad / s), _log(d2 / s), _log(d4 / s),
                 _log(_max_filter(d2, 3) / s), _log(_max_filter(d2, 6) / s),
                 np.full((n, n), math.log10(self.im_e)),
                 np.full((n, n), math.log10(2.0 * math.pi / n)), _log(phic)]
        return np.stack([x.ravel() for x in feats], 1).astype(np.float32)

    def active_cells(self, level: int, phi_l: np.ndarray) -> np.ndarray:
        n = self.n_at(level)
        B = self.bump(n, n)
        return (B.T @ (phi_l > 0).astype(np.float64) @ B) > 0

    # -- the driver ---------------------------------------------------------
    def integrate(self, mode: str = "probe", theta: float = 1e-8, predict=None,
                  margin: float = 0.5, record=None):
        """Run the refinement.  theta: indicator threshold relative to |I|.
        predict(features) -> predicted log10(e/|I|) for mode "learned".
        record(level, features, log10(e/|I|), active) collects training data
        (probe mode).  Returns dict(integral, n_evals, levels, flagged)."""
        if mode not in ("uniform", "probe", "learned"):
            raise ValueError(mode)
        if mode == "learned" and predict is None:
            raise ValueError("learned mode needs predict")
        self.flags = []
        n = self.n_at(0)
        self.ensure(0, np.ones((n, n), bool))
        scale = abs(np.mean(self.values(0)))
        total = 0.0 + 0.0j
        flagged = []
        level = 0
        phi = np.ones((n, n))
        while True:
            n = self.n_at(level)
            self.ensure(level, phi > 0)
            f = np.where(phi > 0, self.values(level), 0.0)
            if level == self.lmax:
                total += np.sum(f * phi) / n ** 2
                break
            act = self.active_cells(level, phi)
            if mode == "uniform":
                U = act.copy()
            elif mode == "probe":
                e = self.indicator(level, phi)
                le = _log(e / scale)
                if record is not None:
                    record(level, self.cell_features(level, phi, scale), le.ravel(), act.ravel())
                U = act & (le > math.log10(theta))
            else:
                pred = np.asarray(predict(self.cell_features(level, phi, scale)),
                                  dtype=np.float64).reshape(n, n)
                U = act & (pred > math.log10(theta) - margin)
            self.flags.append(U)
            flagged.append(int(U.sum()))
            B = self.bump(n, n)
            W = B @ U.astype(np.float64) @ B.T
            total += np.sum(f * phi * (1.0 - W)) / n ** 2
            if not U.any():
                break
            phi = self.window(level + 1, level + 1)
            level += 1
        return {"integral": complex(total), "n_evals": self.n_evals, "levels": level,
                "flagged": flagged}


def uniform_levels(evaluate, n0: int, levels: int, tol: float | None = None,
                   return_grid: bool = False):
    """Plain uniform refinement: the trapezoid rule on the nested N0 2^l grids,
    l = 0..levels, each point evaluated once.  Returns a list of
    (integral, evaluations so far), and the finest grid of values with
    return_grid.  With tol, stops at the first level that agrees with the
    previous one to tol relative (the convergence test)."""
    M = n0 << levels
    if M > MAX_FINEST:
        raise ValueError(f"finest lattice {M} exceeds {MAX_FINEST} points per side")
    grid = np.full((M, M), np.nan + 1j * np.nan, dtype=np.complex128)
    out, n_evals = [], 0
    for l in range(levels + 1):
        s = M // (n0 << l)
        need = np.isnan(grid[::s, ::s].real)
        ix, iy = np.nonzero(need)
        k = -math.pi + 2.0 * math.pi * np.stack([ix, iy], 1).astype(np.float64) / (n0 << l)
        grid[ix * s, iy * s] = np.asarray(evaluate(k), dtype=np.complex128)
        n_evals += len(ix)
        out.append((complex(np.mean(grid[::s, ::s])), n_evals))
        if tol is not None and l > 0 and abs(out[-1][0] - out[-2][0]) <= tol * abs(out[-1][0]):
            break
    return (out, grid) if return_grid else out
