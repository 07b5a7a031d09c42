"""Fermi-Dirac occupations by bisection on the Fermi level (host FP64).

Follows ``src/Eigensolvers/occupations.py:5-56`` and ``FermiDirac.py:3-31``:
kT = Temp * 6.33327186e-06 applied to Hartree eigenvalues (the constant is k_B
in Ry/K, kept verbatim), per-spin occupations in [0, 1], bisection on
[min(lam)-1, lam[ceil(Ne/2)]+1] until |2 sum f - Ne| <= tol (<= 200 steps).
The eigenvalue list has at most a few hundred entries, so this runs on the host
with no device synchronisation per step (the reference's GPU twin syncs on
every bisection step, occupations_gpu.py:26-58).
"""
from __future__ import annotations

import math

import numpy as np

KT_PER_K = 6.33327186e-06


def fermi_dirac(lam, ef, temp, nelec):
    kt = temp * KT_PER_K
    logits = (lam - ef) / kt
    e = np.exp(-np.abs(logits))
    occ = np.where(logits >= 0, e / (1.0 + e), 1.0 / (1.0 + e))
    return float(np.sum(occ) - nelec / 2), occ


def occupations(lam, temp, nelec, tol=1e-6):
    lam = np.asarray(lam, dtype=np.float64).reshape(-1)
    a = float(np.min(lam)) - 1
    fa, _ = fermi_dirac(lam, a, temp, nelec)
    lmax = min(max(1, int(math.ceil(nelec / 2)) + 1), len(lam))
    b = float(lam[lmax - 1]) + 1
    fb, _ = fermi_dirac(lam, b, temp, nelec)
    c = (b + a) / 2
    fc, occ = fermi_dirac(lam, c, temp, nelec)
    error = 2 * np.sum(occ) - nelec
    if fa * fb > tol:
        print("In bisect - fa*fb > 0")
        return b, np.ones(len(lam))
    its = 0
    while abs(error) > tol and its < 200:
        its += 1
        c = (b + a) / 2
        fc, occ = fermi_dirac(lam, c, temp, nelec)
        error = 2 * np.sum(occ) - nelec
        if fc * fb < 0:
            a = c
        else:
            b = c
            fb = fc
    return c, occ
