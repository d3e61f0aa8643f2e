"""Observables of the BZ-integrated tau and their error measures.

tr G(E)  := sum over interface sites of tr tau_BZ,site(E)
DOS(E)   := -(1/pi) Im tr G(E)   (the spectral form of
            skkr_kpar_v6/nrsp-scf/locquant.f:96-103 with F_l = 1, G_l = 0:
            the synthetic system carries no radial wavefunctions)
rel_err  := |A - A_ref|_F / |A_ref|_F   (the relative Frobenius convention of
            compmat, skkr_kpar_v6/main/matops.f:51-80)
integrated G := sum_E w_E tr G(E) over the contour weights of zmesh.
"""
from __future__ import annotations

import numpy as np


def tr_g(tau_bz: np.ndarray) -> complex:
    return complex(np.einsum("sll->", tau_bz))


def site_traces(tau_bz: np.ndarray) -> np.ndarray:
    return np.einsum("sll->s", tau_bz)


def dos(tau_bz: np.ndarray) -> float:
    return float(-tr_g(tau_bz).imag / np.pi)


def rel_err(a, a_ref) -> float:
    a = np.asarray(a)
    a_ref = np.asarray(a_ref)
    den = np.linalg.norm(a_ref)
    return float(np.linalg.norm(a - a_ref) / den) if den > 0 else float(np.linalg.norm(a))


def compmat(x, y, tol=1e-12) -> bool:
    """matops.f:51-80: x equals y when |x-y|_F < tol |x|_F."""
    return bool(np.linalg.norm(np.asarray(x) - np.asarray(y)) < tol * np.linalg.norm(x))


def integrated_g(tr_g_per_energy, weights) -> complex:
    return complex(np.sum(np.asarray(weights) * np.asarray(tr_g_per_energy)))


def error_report(tau, tau_ref):
    """Errors of one energy's result against a reference result."""
    return {"tau_rel_err": rel_err(tau, tau_ref),
            "trG_rel_err": abs(tr_g(tau) - tr_g(tau_ref)) / abs(tr_g(tau_ref)),
            "dos_rel_err": abs(dos(tau) - dos(tau_ref)) / abs(dos(tau_ref))}
