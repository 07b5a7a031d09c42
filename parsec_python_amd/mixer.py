"""Potential mixers on the device: msecant1 (the reference's default,
multisecant Type-I, Broyden's first method), msecant2 (Type-II, Broyden's
second method) and msecant3 (the hybrid), group size 1, mix 0.5, restart
factor 0.1.

Follows ``src/Mixer/msecant1.py:40-175`` with the defaults of
``src/Mixer/includemix.py:21-49`` (EN_like 0, group_size 1) and the 1x1
quasi-inverse of ``src/Mixer/qinv.py``.  The stored columns DX, DF live on the
device (FP64, rows of a (capacity, n) tensor); dot products and linear
combinations are pc_dots / pc_lincomb calls, and only the handful of scalar
coefficients per call come back to the host.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .device import F64
from .hamiltonian import VecOps


class DeviceMsecant1:
    def __init__(self, n: int, capacity: int, mix: float = 0.5, restart_factor: float = 0.1,
                 tol: float = float(np.finfo(float).eps), device_scalars: bool = True):
        # device_scalars: the dot products feed the updates on the device
        # (pc_lincomb_dev), one host readback per call (the restart test)
        # instead of seven; the same arithmetic, so the same bits
        self.device_scalars = device_scalars
        self.n = n
        self.cap = capacity
        self.mix = mix
        self.restart_factor = restart_factor
        self.tol = tol
        self.DX = torch.empty((capacity, n), dtype=F64, device="cuda")
        self.DF = torch.empty((capacity, n), dtype=F64, device="cuda")
        self.tmp = torch.empty((2, n), dtype=F64, device="cuda")
        self.m = 0
        self.vec = VecOps(n)

    def reset(self):
        self.m = 0

    def _dots(self, A, m, v):
        return self.vec.dots(A, m, v).cpu().numpy()

    def __call__(self, x1: torch.Tensor, f1: torch.Tensor, out: torch.Tensor) -> int:
        """out <- new iterate; returns the number of secant pairs used (reference ``m``)."""
        mix = self.mix
        m = self.m
        DX, DF, vec = self.DX, self.DF, self.vec
        if m == 0:
            DX[0].copy_(x1)
            DF[0].copy_(f1)
            self.m = 1
            out.copy_(x1)
            vec.lincomb(out, 1.0, f1, mix)
            return 0
        if m >= 2:
            if self.device_scalars:
                nn = torch.cat([vec.dots(DF[m - 1:m], 1, DF[m - 1]), vec.dots(f1.view(1, -1), 1, f1)]).cpu().numpy()
                nprev, nf = math.sqrt(float(nn[0])), math.sqrt(float(nn[1]))
            else:
                nprev = math.sqrt(float(self._dots(DF[m - 1:m], 1, DF[m - 1])[0]))
                nf = math.sqrt(float(self._dots(f1.view(1, -1), 1, f1)[0]))
            if nprev < self.restart_factor * nf:
                DX[0].copy_(DX[m - 1])
                DF[0].copy_(DF[m - 1])
                self.m = 1
                out.copy_(DX[0])
                vec.lincomb(out, 1.0, DF[0], mix)
                return 0
        if m + 1 > self.cap:
            raise RuntimeError("mixer capacity exceeded")
        if self.device_scalars:
            return self._update_on_device(x1, f1, out)
        ng = m - 1
        last_x, last_f = DX[m - 1], DF[m - 1]
        vec.lincomb(last_x, -1.0, x1, 1.0)          # dx = x1 - x_prev
        vec.lincomb(last_f, -1.0, f1, 1.0)          # df = f1 - f_prev
        dx = self.tmp[0]
        dx.copy_(last_x)
        if ng > 0:
            a = self._dots(DF, ng, last_f)
            vec.lincomb(last_x, 1.0, last_f, mix, A=DX, coef=-a)
        else:
            vec.lincomb(last_x, 1.0, last_f, mix)
        out.copy_(x1)
        if ng > 0:
            b = self._dots(DF, ng, f1)
            vec.lincomb(out, 1.0, f1, mix, A=DX, coef=-b)
        else:
            vec.lincomb(out, 1.0, f1, mix)
        nvec = self.tmp[1]
        nvec.zero_()
        if ng > 0:
            c = self._dots(DX, ng, dx)
            vec.lincomb(nvec, 0.0, dx, -mix, A=DF, coef=c)
        else:
            vec.lincomb(nvec, 0.0, dx, -mix)
        mval = float(self._dots(nvec.view(1, -1), 1, last_f)[0])
        cinv = 0.0 if abs(mval) < self.tol * abs(mval) else 1.0 / mval
        last_f.copy_(nvec)
        vec.lincomb(last_f, cinv)
        d = float(self._dots(last_f.view(1, -1), 1, f1)[0])
        vec.lincomb(out, 1.0, last_x, -d)
        DX[m].copy_(x1)
        DF[m].copy_(f1)
        self.m = m + 1
        return m

    def _update_on_device(self, x1: torch.Tensor, f1: torch.Tensor, out: torch.Tensor) -> int:
        """The secant update of __call__ with every dot product kept on the
        device (src/Mixer/msecant1.py:40-175): the same steps and arithmetic."""
        mix, m = self.mix, self.m
        DX, DF, vec = self.DX, self.DF, self.vec
        ng = m - 1
        last_x, last_f = DX[m - 1], DF[m - 1]
        vec.lincomb(last_x, -1.0, x1, 1.0)          # dx = x1 - x_prev
        vec.lincomb(last_f, -1.0, f1, 1.0)          # df = f1 - f_prev
        dx = self.tmp[0]
        dx.copy_(last_x)
        if ng > 0:
            vec.lincomb_dev(last_x, 1.0, last_f, mix, A=DX, coef=vec.dots(DF, ng, last_f), coef_scale=-1.0)
        else:
            vec.lincomb(last_x, 1.0, last_f, mix)
        out.copy_(x1)
        if ng > 0:
            vec.lincomb_dev(out, 1.0, f1, mix, A=DX, coef=vec.dots(DF, ng, f1), coef_scale=-1.0)
        else:
            vec.lincomb(out, 1.0, f1, mix)
        nvec = self.tmp[1]
        nvec.zero_()
        if ng > 0:
            vec.lincomb_dev(nvec, 0.0, dx, -mix, A=DF, coef=vec.dots(DX, ng, dx))
        else:
            vec.lincomb(nvec, 0.0, dx, -mix)
        mval = vec.dots(nvec.view(1, -1), 1, last_f)
        last_f.copy_(nvec)
        vec.lincomb_dev(last_f, alpha_recip=mval, alpha_tol=self.tol)   # last_f = nvec / mval
        d = vec.dots(last_f.view(1, -1), 1, f1)
        vec.lincomb_dev(out, 1.0, last_x, beta_dev=d, beta_scale=-1.0)
        DX[m].copy_(x1)
        DF[m].copy_(f1)
        self.m = m + 1
        return m


def _column_scale(vec: VecOps, dots, v: torch.Tensor, tol: float) -> float:
    """1/|v|^2, the quasi-inverse scale of one column (src/Mixer/qinv.py:29-62
    on an N x 1 matrix), or 0 when |v| < tol * max|v_i|."""
    nn = float(dots(v.view(1, -1), 1, v)[0])
    r = math.sqrt(max(nn, 0.0))
    vmax = float(v.abs().max().item())
    if r == 0.0 or r < tol * vmax:
        return 0.0
    return 1.0 / nn


class DeviceMsecant2(DeviceMsecant1):
    """Multisecant Type-II (Broyden's second method), group size 1, EN_like 0:
    ``src/Mixer/msecant2.py:24-92``.  A restart keeps (x1, f1) and sets the
    stage to 2; the stage then alternates 2 -> 1 -> 2 as in the reference
    (stage 2 stores no new pair, stage 1 applies the stored pairs unchanged)."""

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self.stage = 0

    def reset(self):
        self.m = 0
        self.stage = 0

    def __call__(self, x1, f1, out) -> int:
        mix, m = self.mix, self.m
        DX, DF, vec = self.DX, self.DF, self.vec
        if m == 0:
            DX[0].copy_(x1)
            DF[0].copy_(f1)
            self.m, self.stage = 1, 0
            out.copy_(x1)
            vec.lincomb(out, 1.0, f1, mix)
            return 0
        if self.stage != 1 and m >= 2:
            nprev = math.sqrt(float(self._dots(DF[m - 1:m], 1, DF[m - 1])[0]))
            nf = math.sqrt(float(self._dots(f1.view(1, -1), 1, f1)[0]))
            if nprev < self.restart_factor * nf:
                DX[0].copy_(x1)
                DF[0].copy_(f1)
                self.m, self.stage = 1, 2
                out.copy_(x1)
                vec.lincomb(out, 1.0, f1, mix)
                return 0
        if m + 1 > self.cap:
            raise RuntimeError("mixer capacity exceeded")
        ng = m - 1
        last_x, last_f = DX[m - 1], DF[m - 1]
        if self.stage != 1:
            vec.lincomb(last_x, -1.0, x1, 1.0)          # dx = x1 - x_prev
            vec.lincomb(last_f, -1.0, f1, 1.0)          # df = f1 - f_prev
            if ng > 0:
                a = self._dots(DF, ng, last_f)
                vec.lincomb(last_x, 1.0, last_f, mix, A=DX, coef=-a)
            else:
                vec.lincomb(last_x, 1.0, last_f, mix)
        out.copy_(x1)
        if ng > 0:
            b = self._dots(DF, ng, f1)
            vec.lincomb(out, 1.0, f1, mix, A=DX, coef=-b)
        else:
            vec.lincomb(out, 1.0, f1, mix)
        if self.stage != 1:
            vec.lincomb(last_f, _column_scale(vec, self._dots, last_f, self.tol))
        d = float(self._dots(last_f.view(1, -1), 1, f1)[0])
        vec.lincomb(out, 1.0, last_x, -d)
        if self.stage != 2:
            DX[m].copy_(x1)
            DF[m].copy_(f1)
            self.m = m + 1
        self.stage = {1: 2, 2: 1}.get(self.stage, self.stage)
        return m


class DeviceMsecant3(DeviceMsecant1):
    """Hybrid multisecant, group size 1, EN_like 0, preferred type 1:
    ``src/Mixer/msecant3.py:45-150``.  Each update is Broyden-I or Broyden-II,
    Type I when |df.df_prev| |N.df| > |df.df| |dx.dx_prev| (N = -mix dx +
    sum_i DF_i (DX_i . dx)), Type II when smaller, Type I on a tie and at the
    first update after a (re)start; a restart drops the history."""

    def __init__(self, n: int, capacity: int, *a, preferred_type: int = 1, **k):
        super().__init__(n, capacity, *a, **k)
        self.prev = torch.zeros((2, n), dtype=F64, device="cuda")   # dx, df of the previous update
        self.preferred = preferred_type
        self.last_type = None

    def __call__(self, x1, f1, out) -> int:
        mix, m = self.mix, self.m
        DX, DF, vec = self.DX, self.DF, self.vec
        if m >= 2:
            nprev = math.sqrt(float(self._dots(DF[m - 1:m], 1, DF[m - 1])[0]))
            nf = math.sqrt(float(self._dots(f1.view(1, -1), 1, f1)[0]))
            if nprev < self.restart_factor * nf:
                m = self.m = 0
        if m == 0:
            DX[0].copy_(x1)
            DF[0].copy_(f1)
            self.prev.zero_()
            self.m = 1
            out.copy_(x1)
            vec.lincomb(out, 1.0, f1, mix)
            return 0
        if m + 1 > self.cap:
            raise RuntimeError("mixer capacity exceeded")
        ng = m - 1
        last_x, last_f = DX[m - 1], DF[m - 1]
        vec.lincomb(last_x, -1.0, x1, 1.0)          # dx = x1 - x_prev
        vec.lincomb(last_f, -1.0, f1, 1.0)          # df = f1 - f_prev
        dx = self.tmp[0]
        dx.copy_(last_x)
        if ng > 0:
            a = self._dots(DF, ng, last_f)
            vec.lincomb(last_x, 1.0, last_f, mix, A=DX, coef=-a)
        else:
            vec.lincomb(last_x, 1.0, last_f, mix)
        out.copy_(x1)
        if ng > 0:
            b = self._dots(DF, ng, f1)
            vec.lincomb(out, 1.0, f1, mix, A=DX, coef=-b)
        else:
            vec.lincomb(out, 1.0, f1, mix)
        nvec = self.tmp[1]
        nvec.zero_()
        if ng > 0:
            c = self._dots(DX, ng, dx)
            vec.lincomb(nvec, 0.0, dx, -mix, A=DF, coef=c)
        else:
            vec.lincomb(nvec, 0.0, dx, -mix)
        if ng > 0:
            d_fp = float(self._dots(self.prev[1:2], 1, last_f)[0])
            d_nf = float(self._dots(nvec.view(1, -1), 1, last_f)[0])
            d_ff = float(self._dots(last_f.view(1, -1), 1, last_f)[0])
            d_xp = float(self._dots(self.prev[0:1], 1, dx)[0])
            lhs, rhs = abs(d_fp) * abs(d_nf), abs(d_ff) * abs(d_xp)
            typ = 1 if lhs > rhs else (2 if lhs < rhs else self.preferred)
        else:
            typ = self.preferred
        self.prev[0].copy_(dx)
        self.prev[1].copy_(last_f)
        if typ == 1:
            mval = float(self._dots(nvec.view(1, -1), 1, last_f)[0])
            cinv = 0.0 if abs(mval) < self.tol * abs(mval) else 1.0 / mval
            last_f.copy_(nvec)
            vec.lincomb(last_f, cinv)
        else:
            vec.lincomb(last_f, _column_scale(vec, self._dots, last_f, self.tol))
        d = float(self._dots(last_f.view(1, -1), 1, f1)[0])
        vec.lincomb(out, 1.0, last_x, -d)
        DX[m].copy_(x1)
        DF[m].copy_(f1)
        self.m = m + 1
        self.last_type = typ
        return m


def make_mixer(name: str, n: int, capacity: int):
    """The reference's mixer names (src/Mixer/mixer.py:37-107)."""
    kinds = {"msecant1": DeviceMsecant1, "msecant2": DeviceMsecant2, "msecant3": DeviceMsecant3}
    if name not in kinds:
        raise ValueError(f"Unknown mixer '{name}'")
    return kinds[name](n, capacity)
