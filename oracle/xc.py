"""LDA exchange-correlation, Ceperley-Alder / Perdew-Zunger, Rydberg units.

Vectorised restatement of ``src/V_xc/exc_nspn.py:8-103`` (the reference loops
over points in Python; lines 73-91).  Input rho in e/bohr^3; returns
(v_xc [Ry] per point, E_xc [Ry] = h^3 sum(3/4 rho v_x + rho e_c)).
"""
from __future__ import annotations

import numpy as np

G, B1, B2 = -0.2846, 1.0529, 0.3334
C1, C2, C3, C4, C5 = 0.0622, 0.096, 0.004, 0.0232, 0.0192
THIRD = 1.0 / 3.0
A0 = (4.0 / (9.0 * np.pi)) ** THIRD
TWOVPIA0 = 2.0 / (np.pi * A0)
P75VPI = 0.75 / np.pi


def exc_nspn(h: float, rho: np.ndarray):
    rho = np.asarray(rho, dtype=np.float64)
    v = np.zeros_like(rho)
    pos = rho > 0.0
    rp = rho[pos]
    rs = (P75VPI / rp) ** THIRD
    vx = -TWOVPIA0 / rs
    e = 0.75 * rp * vx
    hi = rs >= 1.0
    ec = np.empty_like(rs)
    vc = np.empty_like(rs)
    sq = np.sqrt(rs[hi])
    ech = G / (1.0 + B1 * sq + B2 * rs[hi])
    ec[hi] = ech
    vc[hi] = (ech ** 2) * (1.0 + 3.5 * B1 * sq * THIRD + 4.0 * B2 * rs[hi] * THIRD) / G
    lo = ~hi
    al = np.log(rs[lo])
    ecl = C1 * al - C2 + (C3 * al - C4) * rs[lo]
    ec[lo] = ecl
    vc[lo] = ecl - (C1 + (C3 * al - C5) * rs[lo]) * THIRD
    v[pos] = vx + vc
    e = e + rp * ec
    return v, float(np.sum(e)) * h ** 3


def exc_nspn_loop(h: float, rho: np.ndarray):
    """The same functional evaluated point by point in a Python loop, in the
    reference's order (``src/V_xc/exc_nspn.py:73-91``, energy accumulated as
    it goes).  Used where the reference's own cost is being timed."""
    rho = np.asarray(rho, dtype=np.float64)
    v = rho.copy()
    exc = 0.0
    sqrt, log = np.sqrt, np.log
    for i in range(v.shape[0]):
        r = v[i]
        v[i] = 0.0
        if r > 0.0:
            rs = (P75VPI / r) ** THIRD
            v[i] = -TWOVPIA0 / rs
            exc += 0.75 * r * v[i]
            if rs >= 1.0:
                sq = sqrt(rs)
                ec = G / (1.0 + B1 * sq + B2 * rs)
                v[i] += (ec ** 2) * (1.0 + 3.5 * B1 * sq * THIRD + 4.0 * B2 * rs * THIRD) / G
            else:
                al = log(rs)
                ec = C1 * al - C2 + (C3 * al - C4) * rs
                v[i] += ec - (C1 + (C3 * al - C5) * rs) * THIRD
            exc += r * ec
    return v, exc * h ** 3
