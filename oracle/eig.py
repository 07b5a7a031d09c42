"""Chebyshev-filtered subspace iteration, Lanczos bounds, occupations (oracle).

``H`` is anything supporting ``H @ X`` (SciPy sparse in practice).  Random
numbers come from an explicit ``numpy.random.Generator`` (the reference is
unseeded, P6): ``rand`` for ``lancz_uppbnd``/``first_filt`` start vectors and
``randn`` for the chefsi1 Lanczos start (chefsi1.py:29).

Follows:
* ``lanczos_bound``     src/Eigensolvers/lanczosForChefsi1.py:5-99 (6 steps)
* ``lancz_uppbnd``      src/Eigensolvers/lancz_uppbnd.py:4-75
* ``ch_filter``         src/Eigensolvers/ch_filter.py:4-81
* ``cheb_filter_slim``  src/Eigensolvers/first_filt.py:78-103
* ``first_filt``        src/Eigensolvers/first_filt.py:6-75
* ``chefsi1``           src/Eigensolvers/chefsi1.py:10-69
* ``fermi_dirac``       src/Eigensolvers/FermiDirac.py:3-31
* ``occupations``       src/Eigensolvers/occupations.py:5-56
"""
from __future__ import annotations

import math

import numpy as np
from scipy.linalg import eigh, qr

KT_PER_K = 6.33327186e-06  # FermiDirac.py:16 (k_B in Ry/K applied to Ha, P5)


def lanczos_bound(H, v, m: int = 6, nev: int = 6) -> float:
    """Upper spectral bound max|theta| + beta_m after m plain Lanczos steps."""
    n = v.shape[0]
    v = v.reshape(-1) / np.linalg.norm(v)
    v1 = v
    v0 = np.zeros(n)
    bet = 0.0
    tmat = np.zeros((m + 1, m + 1))
    bound = None
    for k in range(1, m + 1):
        w = H @ v1 - bet * v0
        alp = float(np.vdot(v1, w))
        w = w - alp * v1
        bet = float(np.linalg.norm(w))
        v0 = v1
        v1 = w / bet
        tmat[k - 1, k - 1] = alp
        tmat[k, k - 1] = bet
        tmat[k - 1, k] = bet
        ntest = min(5 * nev, m)
        if (k >= ntest and k % 10 == 0) or k == m:
            rr = np.sort(eigh(tmat[:k, :k], eigvals_only=True))
            bound = max(np.abs(rr)) + bet
    return float(bound)


def lancz_uppbnd(H, v, k: int = 6):
    """Upper bound and Ritz values from k (6..10) Lanczos steps (lancz_uppbnd.py)."""
    k = min(max(k, 6), 10)
    t = np.zeros((k, k))
    v = v.reshape(-1) / np.linalg.norm(v)
    tol = 2.5e-16
    upperb = np.zeros((3, k))
    f = H @ v
    alpha = float(np.vdot(v, f))
    f = f - alpha * v
    t[0, 0] = alpha
    beta = float(np.linalg.norm(f))
    upperb[:, 0] = alpha + beta
    ritzv = None
    j = 1
    for j in range(2, k + 1):
        if beta > tol:
            v0 = v
            v = f / beta
            f = H @ v
            f = f - beta * v0
            alpha = float(np.vdot(v, f))
            f = f - alpha * v
            t[j - 1, j - 2] = beta
            t[j - 2, j - 1] = beta
            t[j - 1, j - 1] = alpha
        else:
            raise RuntimeError("invariant subspace found in lancz_uppbnd")
        beta = float(np.linalg.norm(f))
        ritzv, x = eigh(t[0:j + 1, 0:j + 1])
        if beta < 1e-2:
            beta = beta * 10
        upperb[0, j - 1] = ritzv[-1] + beta
        upperb[1, j - 1] = ritzv[-1] + abs(x[-1, -1]) * beta
        upperb[2, j - 1] = ritzv[-1] + max(abs(x[-1, :])) * beta
    u = (upperb[0, j - 1] + upperb[1, j - 1]) / 2
    u = (upperb[2, j - 1] + u) / 2
    return float(u), ritzv


def ch_filter(H, x, deg, lam1, low, high):
    e = (high - low) / 2
    c = (high + low) / 2
    sigma1 = e / (lam1 - c)
    sigma = sigma1
    y = (H @ x - c * x) * (sigma1 / e)
    tau = 2.0 / sigma1
    inv_e = 1.0 / e
    for _ in range(2, deg + 1):
        sigma_new = 1.0 / (tau - sigma)
        t1 = 2.0 * sigma_new * inv_e
        t2 = sigma * sigma_new
        ynew = (H @ y - c * y) * t1 - t2 * x
        x = y
        y = ynew
        sigma = sigma_new
    return y


def cheb_filter_slim(H, x, polm, low, high):
    e = (high - low) / 2
    center = (high + low) / 2
    y = H @ x
    y = (-y + center * x) / e
    for _ in range(2, polm + 1):
        ynew = H @ y
        ynew = (-ynew + center * y) * 2 / e - x
        x = y
        y = ynew
    return y


def rayleigh_ritz(H, W):
    """QR, G = W^T H W (symmetrised from the upper triangle), eigh, rotate."""
    W, _ = qr(W, mode="economic", pivoting=False)
    V = H @ W
    G = np.triu(V.T @ W)
    G = G + np.triu(G, 1).T
    ritzv_raw, Q = eigh(G)
    order = np.argsort(ritzv_raw)
    return W @ Q[:, order], ritzv_raw[order]


def first_filt(nev, H, polm, rng: np.random.Generator):
    """Returns (W, ritzv, info); ``nev`` is the block width (nev+15 at the caller)."""
    energ_tol = 0.08
    if polm < 10 or polm > 15:
        polm = 10
    max_iter = max(min(int(60 / polm), 5), 3)
    n = H.shape[0]
    upperb, ritzv = lancz_uppbnd(H, rng.random((n, 1)))
    lowerb = 0.8 * ritzv[0] + 0.2 * ritzv[-2]
    W = rng.random((n, nev))
    tr0 = None
    it_done = 0
    for it in range(1, max_iter + 1):
        it_done = it
        W = cheb_filter_slim(H, W, polm, lowerb, upperb)
        W, ritzv = rayleigh_ritz(H, W)
        lowerb = ritzv[-1]
        if it == 1:
            tr0 = np.sum(ritzv[:nev])
        else:
            tr1 = np.sum(ritzv[:nev])
            if abs(tr1 - tr0) <= energ_tol * abs(tr0):
                break
            tr0 = tr1
    return W, ritzv, {"outer_iterations": it_done, "upperb": upperb}


def chefsi1(Vin, ritzv, deg, nev, H, rng: np.random.Generator):
    n, n2 = Vin.shape
    if nev >= n2:
        raise ValueError("Vin should have more than nev columns")
    upperb = lanczos_bound(H, rng.standard_normal((n, 1)), 6, 6)
    lowerb = float(np.max(ritzv))
    lam1 = float(np.min(ritzv))
    if lowerb > upperb:
        raise ValueError("Bounds are incorrect")
    W = ch_filter(H, Vin, deg, lam1, lowerb, upperb)
    W, ritzv = rayleigh_ritz(H, W)
    return W, ritzv


def fermi_dirac(lam, ef, temp, nelec):
    kt = temp * KT_PER_K
    logits = (lam - ef) / kt
    e = np.exp(-np.abs(logits))
    occ = np.where(logits >= 0, e / (1.0 + e), 1.0 / (1.0 + e))
    return float(np.sum(occ) - nelec / 2), occ


def occupations(lam, temp, nelec, tol=1e-6):
    lam = np.asarray(lam, dtype=np.float64)
    its = 0
    a = float(np.min(lam)) - 1
    fa, _ = fermi_dirac(lam, a, temp, nelec)
    lmax = min(max(1, int(math.ceil(nelec / 2)) + 1), len(lam))
    b = float(lam[lmax - 1]) + 1
    fb, _ = fermi_dirac(lam, b, temp, nelec)
    c = (b + a) / 2
    fc, occ = fermi_dirac(lam, c, temp, nelec)
    error = 2 * np.sum(occ) - nelec
    if fa * fb > tol:
        return b, np.ones(len(lam))
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
