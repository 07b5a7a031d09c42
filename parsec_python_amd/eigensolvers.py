"""Chebyshev-filtered subspace iteration on the device.

Mirrors the reference functions (same arguments, same algorithms):
* ``lanczos_bound``   -- ``src/Eigensolvers/lanczosForChefsi1.py:5-99`` (6 steps,
                         bound = max|theta| + beta_6)
* ``lancz_uppbnd``    -- ``src/Eigensolvers/lancz_uppbnd.py:4-75``
* ``ch_filter``       -- ``src/Eigensolvers/ch_filter.py:4-81`` (scaled filter)
* ``cheb_filter_slim``-- ``src/Eigensolvers/first_filt.py:78-103``
* ``first_filt``      -- ``src/Eigensolvers/first_filt.py:6-75``
* ``chefsi1``         -- ``src/Eigensolvers/chefsi1.py:10-69``

Orthonormalisation + Rayleigh-Ritz (the reference's QR, G = W^T H W, eigh,
W Q) is done as S = Y^T Y, G_y = Y^T (H Y) (MFMA Gram), FP64 Cholesky /
triangular inverse / symmetric eigensolver on the GPU, and one rotation
Y (R^-1 Q): the same Ritz pairs.

Random start vectors come from the counter-based generator of pc_rand_* with
an explicit seed (the reference is unseeded).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib
from .device import F64, Scratch, ptr, stream
from .hamiltonian import BlockLinalg, Hamiltonian, VecOps


class SeedStream:
    """Deterministic per-call seeds."""

    def __init__(self, seed: int):
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.count = 0

    def next(self) -> int:
        self.count += 1
        return (self.seed * 0x100000001B3 + self.count * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF


def rand_block(out: torch.Tensor, n: int, k: int, seed: int, normal: bool) -> None:
    suffix = "f64" if out.dtype == F64 else "f32"
    _lib.call(f"pc_rand_{suffix}", n, k, out.shape[1], seed, int(normal), ptr(out), stream())


class Eigensolver:
    def __init__(self, H: Hamiltonian, kmax: int, seed: int = 0):
        self.H = H
        self.n = H.n
        self.ld = H.ld
        self.kmax = kmax
        self.seeds = SeedStream(seed)
        self.vec = VecOps(self.n)
        self.la = BlockLinalg(self.n, kmax)
        self.bufs = [H.new_block(kmax) for _ in range(3)]
        self.lz = torch.empty((3, self.ld), dtype=F64, device="cuda")  # Lanczos vectors (FP64)
        nbl = C.c_int()
        _lib.call("pc_lanczos_nblocks", self.n, C.byref(nbl))
        self.lz_nblocks = nbl.value
        self.scratch = Scratch()
        self.S = torch.empty((kmax, kmax), dtype=F64, device="cuda")
        self.G = torch.empty((kmax, kmax), dtype=F64, device="cuda")
        self.stats = {"filter_steps": 0, "rr": 0}

    # -------------------------------------------------------------- Lanczos
    def _h64_step(self, y, x, out, c):
        """out = H y + c x on single FP64 vectors (the FP64 potential is always kept)."""
        self.H.step(y, x, out, 1, 1.0, 0.0, c)

    def _dot(self, a, b) -> float:
        return float(self.vec.dots(a, 1, b).item())

    def lanczos_bound(self, m: int = 6, nev: int = 6) -> float:
        """Upper bound of the spectrum (lanczosForChefsi1.py:5-99): m Lanczos
        steps from a random start, max |Ritz value| of T_m plus the last beta.
        alpha and beta stay on the device (pc_lanczos_step); the host reads them
        once at the end."""
        v0, v1, w = self.lz[0:1], self.lz[1:2], self.lz[2:3]
        rand_block(v1, self.n, 1, self.seeds.next(), normal=True)
        coef = self.scratch.get("lz_coef", 2 * m + 2)
        part = self.scratch.get("lz_part", 2 * self.lz_nblocks)
        _lib.call("pc_lanczos_start", self.n, ptr(v1), ptr(part), stream())
        v0.zero_()
        for k in range(m):
            self._h64_step(v1, v0, w, 0.0)               # w = H v1
            _lib.call("pc_lanczos_step", self.n, ptr(v0), ptr(v1), ptr(w), ptr(coef), k, ptr(part), stream())
            v0, v1, w = v1, w, v0
        c = coef.cpu().numpy()
        alpha, beta = c[0:2 * m:2], c[1:2 * m:2]
        tmat = np.diag(alpha) + np.diag(beta[:m - 1], 1) + np.diag(beta[:m - 1], -1)
        rr = np.linalg.eigvalsh(tmat)
        bound = float(np.max(np.abs(rr)) + beta[m - 1])
        if not np.isfinite(bound):
            raise _lib.ParsecHipError("lanczos_bound: non-finite bound (invariant subspace)")
        return bound

    def lancz_uppbnd(self, k: int = 6):
        k = min(max(k, 6), 10)
        v, v0, f = self.lz[0:1], self.lz[1:2], self.lz[2:3]
        rand_block(v, self.n, 1, self.seeds.next(), normal=False)
        nrm = math.sqrt(self._dot(v, v))
        self.vec.lincomb(v, alpha=1.0 / nrm)
        t = np.zeros((k, k))
        upperb = np.zeros((3, k))
        self._h64_step(v, None, f, 0.0)
        alpha = self._dot(v, f)
        self.vec.lincomb(f, alpha=1.0, x=v, beta=-alpha)
        t[0, 0] = alpha
        beta = math.sqrt(self._dot(f, f))
        upperb[:, 0] = alpha + beta
        ritzv = None
        j = 1
        for j in range(2, k + 1):
            if not beta > 2.5e-16:
                raise _lib.ParsecHipError("lancz_uppbnd: invariant subspace found")
            # v0 <- v ; v <- f / beta ; f <- H v - beta v0
            v0, v, f = v, f, v0
            self.vec.lincomb(v, alpha=1.0 / beta)
            self._h64_step(v, v0, f, -beta)
            alpha = self._dot(v, f)
            self.vec.lincomb(f, alpha=1.0, x=v, beta=-alpha)
            t[j - 1, j - 2] = beta
            t[j - 2, j - 1] = beta
            t[j - 1, j - 1] = alpha
            beta = math.sqrt(self._dot(f, f))
            sub = t[0:j + 1, 0:j + 1]
            ritzv, x = np.linalg.eigh(sub)
            if beta < 1e-2:
                beta = beta * 10
            upperb[0, j - 1] = ritzv[-1] + beta
            upperb[1, j - 1] = ritzv[-1] + abs(x[-1, -1]) * beta
            upperb[2, j - 1] = ritzv[-1] + max(abs(x[-1, :])) * beta
        u = (upperb[0, j - 1] + upperb[1, j - 1]) / 2
        u = (upperb[2, j - 1] + u) / 2
        return float(u), ritzv

    # -------------------------------------------------------------- filters
    def ch_filter(self, X: torch.Tensor, k: int, deg: int, lam1: float, low: float, high: float) -> torch.Tensor:
        """Scaled Chebyshev filter of X[:k]; returns the buffer holding the result.
        X's buffer is reused (it must be one of self.bufs)."""
        e = (high - low) / 2
        c = (high + low) / 2
        sigma1 = e / (lam1 - c)
        sigma = sigma1
        tau = 2.0 / sigma1
        Xb = X
        Yb = self._other(X)
        # Y1 = (H X - c X) * sigma1/e
        s = sigma1 / e
        self.H.step(Xb, None, Yb, k, s, -c * s, 0.0)
        for _ in range(2, deg + 1):
            sigma_new = 1.0 / (tau - sigma)
            t1 = 2.0 * sigma_new / e
            t2 = sigma * sigma_new
            # Ynew = (H Y - c Y) t1 - t2 X, written over X
            self.H.step(Yb, Xb, Xb, k, t1, -c * t1, -t2)
            Xb, Yb = Yb, Xb
            sigma = sigma_new
        self.stats["filter_steps"] += deg
        return Yb

    def cheb_filter_slim(self, X: torch.Tensor, k: int, polm: int, low: float, high: float) -> torch.Tensor:
        e = (high - low) / 2
        center = (high + low) / 2
        Xb = X
        Yb = self._other(X)
        # y = (-H x + center x)/e
        self.H.step(Xb, None, Yb, k, -1.0 / e, center / e, 0.0)
        for _ in range(2, polm + 1):
            # ynew = (-H y + center y) * 2/e - x
            self.H.step(Yb, Xb, Xb, k, -2.0 / e, 2.0 * center / e, -1.0)
            Xb, Yb = Yb, Xb
        self.stats["filter_steps"] += polm
        return Yb

    def _other(self, X, avoid=None):
        for b in self.bufs:
            if b.data_ptr() != X.data_ptr() and (avoid is None or b.data_ptr() != avoid.data_ptr()):
                return b
        raise RuntimeError("no free block buffer")

    # -------------------------------------------------------------- Rayleigh-Ritz
    def rayleigh_ritz(self, Y: torch.Tensor, k: int):
        HY = self._other(Y)
        self.H.apply(Y, HY, k)
        S = self.S.view(-1)[:k * k]
        G = self.G.view(-1)[:k * k]
        self.la.gram(Y, Y, k, k, S, sym=True)
        # G = Y^T (H Y) is symmetric: only the upper tiles are computed and mirrored
        # (the dense stage symmetrises G again after the Cholesky transform)
        self.la.gram(Y, HY, k, k, G, sym=True)
        T, lam = self.la.dense_rr(S, G, k)
        W = HY  # HY is no longer needed
        self.la.rotate(Y, T, k, k, W)
        self.stats["rr"] += 1
        return W, lam

    # -------------------------------------------------------------- drivers
    def first_filt(self, k: int, polm: int = 10, start_weight: torch.Tensor | None = None):
        """first_filt.py: random start block, slim Chebyshev filter + Rayleigh-Ritz
        passes until the Ritz-value trace settles.  ``start_weight`` (per point,
        optional) multiplies the random start vectors point-wise."""
        energ_tol = 0.08
        if polm < 10 or polm > 15:
            polm = 10
        max_iter = max(min(int(60 / polm), 5), 3)
        upperb, ritzv = self.lancz_uppbnd()
        lowerb = 0.8 * ritzv[0] + 0.2 * ritzv[-2]
        W = self.bufs[0]
        rand_block(W, self.n, k, self.seeds.next(), normal=False)
        if start_weight is not None:
            W[:k, :self.n] *= start_weight.to(W.dtype)
        tr0 = None
        outer = 0
        for it in range(1, max_iter + 1):
            outer = it
            Y = self.cheb_filter_slim(W, k, polm, lowerb, upperb)
            W, ritzv = self.rayleigh_ritz(Y, k)
            lowerb = ritzv[-1]
            if it == 1:
                tr0 = np.sum(ritzv[:k])
            else:
                tr1 = np.sum(ritzv[:k])
                if abs(tr1 - tr0) <= energ_tol * abs(tr0):
                    break
                tr0 = tr1
        self.last_first_filt = {"outer_iterations": outer, "upperb": upperb}
        return W, ritzv

    def chefsi1(self, W: torch.Tensor, ritzv: np.ndarray, deg: int, nev: int):
        k = len(ritzv)
        if nev >= k:
            raise ValueError("Vin should have more than nev columns")
        upperb = self.lanczos_bound(6, 6)
        lowerb = float(np.max(ritzv))
        lam1 = float(np.min(ritzv))
        if lowerb > upperb:
            raise ValueError("Bounds are incorrect")
        self.last_bounds = (lam1, lowerb, upperb)
        Y = self.ch_filter(W, k, deg, lam1, lowerb, upperb)
        return self.rayleigh_ritz(Y, k)

    def close(self):
        self.la.close()
