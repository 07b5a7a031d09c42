"""Chebyshev-filtered subspace iteration on the device.

Mirrors the reference functions (same arguments, same algorithms):
* ``lanczos_bound``   -- ``src/Eigensolvers/lanczosForChefsi1.py:5-99`` (6 steps,
                         bound = max|theta| + beta_6)
* ``lancz_uppbnd``    -- ``src/Eigensolvers/lancz_uppbnd.py:4-75``
* ``ch_filter``       -- ``src/Eigensolvers/ch_filter.py:4-81`` (scaled filter)
* ``cheb_filter_slim``-- ``src/Eigensolvers/first_filt.py:78-103``
* ``first_filt``      -- ``src/Eigensolvers/first_filt.py:6-75``
* ``chefsi1``         -- ``src/Eigensolvers/chefsi1.py:10-69``
* ``lanczos``         -- ``src/Eigensolvers/lanczos.py:9-223`` (diagmeth 0/1)
* ``chsubsp``         -- ``src/Eigensolvers/chsubsp.py:10-91`` (diagmeth 2)

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


def _tridiag(alpha, beta):
    m = len(alpha)
    t = np.diag(np.asarray(alpha, dtype=np.float64))
    if m > 1:
        b = np.asarray(beta[:m - 1], dtype=np.float64)
        t += np.diag(b, 1) + np.diag(b, -1)
    return t


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
    def __init__(self, H: Hamiltonian, kmax: int, seed: int = 0, overlap: bool = True):
        self.H = H
        # Rayleigh-Ritz: S = Y^T Y and its Cholesky factor on a second stream
        # while H Y and G = Y^T H Y are formed on the main one
        self.overlap = overlap
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
        S = self.S.view(-1)[:k * k]
        G = self.G.view(-1)[:k * k]
        if self.overlap:
            # S and its Cholesky factor on a second stream while H Y and G are formed
            self.la.gram_chol_async(Y, k, S)
        else:
            self.la.gram(Y, Y, k, k, S, sym=True)
        self.H.apply(Y, HY, k)
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

    # -------------------------------------------------------------- Lanczos (diagmeth 0/1) and chsubsp (2)
    def _lanczos_run(self, m: int, nev_test: int, conv_test: bool, tol: float):
        """m steps of plain Lanczos (no reorthogonalisation) from a normal random
        start, every vector kept (FP64 rows of VV).  With ``conv_test`` the run
        stops when the sum of the nev_test lowest Ritz values of T_k, checked
        every 10 steps from step min(8 nev_test, m) on, changes by less than tol
        relative (src/Eigensolvers/lanczos.py:64-131; the reference runs this
        test for n <= 35000 only).  Returns (VV, alpha, beta, k)."""
        n, ld = self.n, self.ld
        VV = torch.empty((m + 1, ld), dtype=F64, device="cuda")
        coef = self.scratch.get("lzf_coef", 2 * m + 2)
        part = self.scratch.get("lz_part", 2 * self.lz_nblocks)
        rand_block(VV[0:1], n, 1, self.seeds.next(), normal=True)
        _lib.call("pc_lanczos_start", n, ptr(VV[0:1]), ptr(part), stream())
        zero = self.lz[0:1]
        zero.zero_()
        ntest = min(8 * nev_test, m)
        tr_prev = None
        k = 0
        for j in range(m):
            v0 = VV[j - 1:j] if j > 0 else zero
            v1, w = VV[j:j + 1], VV[j + 1:j + 2]
            self._h64_step(v1, None, w, 0.0)                        # w = H v1
            _lib.call("pc_lanczos_step", n, ptr(v0), ptr(v1), ptr(w), ptr(coef), j, ptr(part), stream())
            k = j + 1
            if conv_test and ((k >= ntest and k % 10 == 0) or k == m):
                c = coef[:2 * k].cpu().numpy()
                rr = np.linalg.eigvalsh(_tridiag(c[0:2 * k:2], c[1:2 * k - 1:2]))
                tr = float(np.sum(rr[:nev_test]))
                if tr_prev is not None and abs(tr - tr_prev) < tol * abs(tr_prev):
                    break
                tr_prev = tr
        c = coef[:2 * k].cpu().numpy()
        return VV, c[0:2 * k:2], c[1:2 * k:2], k

    def _rr_wide(self, Y: torch.Tensor, kk: int, keep: int):
        """Rayleigh-Ritz of the kk-vector FP64 block Y (kk may exceed the
        solver's block width); returns (W (keep, ld) in the solver's dtype, the
        keep lowest Ritz values), or None when Y spans fewer than keep
        independent directions.

        The Ritz vectors of a Lanczos run without reorthogonalisation contain
        near-copies (ghosts), so Y is orthonormalised first, by the eigen-
        decomposition of its scaled Gram matrix D^-1/2 Y^T Y D^-1/2 = U L U^T
        (SVQB): directions with L_i below 1e-10 max L are dropped and
        Y1 = Y D^-1/2 U L^-1/2 has orthonormal columns to rounding.  This plays
        the part of the reference's Householder QR (src/Eigensolvers/
        lanczos.py:194), which turns the dependent columns into normalised
        rounding noise whose Ritz values lie far above the wanted ones."""
        la = BlockLinalg(self.n, kk)
        try:
            S = torch.empty(kk * kk, dtype=F64, device="cuda")
            la.gram(Y, Y, kk, kk, S, sym=True)
            Sm = S.view(kk, kk)
            d = torch.rsqrt(torch.diagonal(Sm).clamp_min(1e-300))
            Sm.mul_(d[:, None] * d[None, :])
            ev = np.zeros(kk)
            _lib.call("pc_dense_syevd", la.dense, kk, ptr(S), kk, ev.ctypes.data_as(C.POINTER(C.c_double)))
            idx = np.nonzero(ev >= 1e-10 * ev.max())[0]
            r = len(idx)
            if r < keep:
                return None
            # column j of U is row j of Sm (column-major storage)
            sel = torch.from_numpy(idx).to("cuda")
            T1 = Sm.index_select(0, sel) * d[None, :] * torch.from_numpy(ev[idx] ** -0.5).to("cuda")[:, None]
            T1 = T1.contiguous().view(-1)
            Y1 = torch.empty((r, self.ld), dtype=F64, device="cuda")
            la.rotate(Y, T1, kk, r, Y1)
            HY = torch.empty_like(Y1)
            dt = self.H.dtype
            self.H.dtype = F64
            try:
                self.H.apply(Y1, HY, r)
            finally:
                self.H.dtype = dt
            S1 = torch.empty(r * r, dtype=F64, device="cuda")
            G1 = torch.empty(r * r, dtype=F64, device="cuda")
            la.gram(Y1, Y1, r, r, S1, sym=True)
            la.gram(Y1, HY, r, r, G1, sym=True)
            T, lam = la.dense_rr(S1, G1, r)
            la.rotate(Y1, T[:r * keep], r, keep, HY)          # the first keep columns of T
        finally:
            la.close()
        W = self.bufs[0]
        if W.dtype == F64:
            W[:keep].copy_(HY[:keep])
        else:
            _lib.call("pc_convert_f64_to_f32", self.n, keep, ptr(HY), HY.shape[1], ptr(W), W.shape[1], stream())
        self.stats["rr"] += 1
        return W, np.asarray(lam[:keep])

    def _lanczos_ritz(self, VV, alpha, beta, k: int, nev: int):
        """Ritz vectors of the Lanczos basis the reference way
        (src/Eigensolvers/lanczos.py:149-209): eigenpairs of T_k, the count of
        distinct Ritz values (within 1e-5) sets mev = min(nev * min(1.8^(k /
        distinct), 3.2), k) vectors V Y, which are orthonormalised and
        Rayleigh-Ritz projected; the nev lowest pairs are returned, with the
        bound max|theta| + beta_k.  Should the mev vectors span fewer than nev
        directions, mev is doubled (up to k)."""
        rr, X1 = np.linalg.eigh(_tridiag(alpha[:k], beta[:k - 1]))
        distinct = [rr[0]]
        for x in rr[1:]:
            if all(abs(x - d) >= 1e-5 for d in distinct[-8:]):
                distinct.append(x)
        mult = min(1.8 ** (len(rr) / len(distinct)), 3.2)
        mev = max(nev, int(min(nev * mult, k)))
        bound = float(np.max(np.abs(rr)) + beta[k - 1])
        while True:
            Q = torch.tensor(np.asfortranarray(X1[:, :mev]).reshape(-1, order="F"), dtype=F64, device="cuda")
            Y = torch.empty((mev, self.ld), dtype=F64, device="cuda")
            la = BlockLinalg(self.n, max(k, mev))
            la.rotate(VV, Q, k, mev, Y)
            la.close()
            out = self._rr_wide(Y, mev, nev)
            if out is not None:
                self.last_lanczos = {"steps": k, "mev": mev, "distinct": len(distinct)}
                return out[0], out[1], bound
            if mev >= k:
                raise _lib.ParsecHipError(f"lanczos: {k} steps span fewer than {nev} independent Ritz vectors")
            mev = min(2 * mev, k)

    def lanczos(self, k: int, m: int, tol: float = 1e-5):
        """The reference's ``lanczos(B, nev, v, m, tol)`` (diagmeth 0 at the
        first SCF iteration, diagmeth 1 every iteration): m Lanczos steps, k
        Ritz pairs (src/Eigensolvers/lanczos.py:9-223)."""
        VV, a, b, kd = self._lanczos_run(m, k, self.n <= 35000, tol)
        W, lam, _ = self._lanczos_ritz(VV, a, b, kd, k)
        return W, lam

    def chsubsp(self, k: int, deg: int):
        """The reference's ``chsubsp(deg, nev, H)`` (diagmeth 2, first SCF
        iteration; src/Eigensolvers/chsubsp.py:10-91): Lanczos with max(3k, 450)
        steps (src/Eigensolvers/lanczosForChsubsp.py:7-193) for the start block,
        its Ritz values and the upper bound; then up to 30 passes of the scaled
        Chebyshev filter on [max ritz, bound] + Rayleigh-Ritz until the sum of the
        k Ritz values changes by less than 5 %."""
        steps = max(3 * k, 450)
        VV, a, b, kd = self._lanczos_run(steps, k, self.n <= 35000, 1e-5)
        W, ritzv, upperb = self._lanczos_ritz(VV, a, b, kd, k)
        del VV
        tr0 = float(np.sum(ritzv))
        outer = 0
        for _ in range(30):
            outer += 1
            lam1, lowerb = float(np.min(ritzv)), float(np.max(ritzv))
            if lowerb > upperb:
                raise ValueError("bounds are wrong")
            Y = self.ch_filter(W, k, deg, lam1, lowerb, upperb)
            W, ritzv = self.rayleigh_ritz(Y, k)
            tr1 = float(np.sum(ritzv[:k]))
            if abs(tr1 - tr0) < 0.05 * abs(tr1):
                break
            tr0 = tr1
        self.last_chsubsp = {"lanczos_steps": kd, "outer_iterations": outer, "upperb": upperb}
        return W, ritzv

    def close(self):
        self.la.close()
