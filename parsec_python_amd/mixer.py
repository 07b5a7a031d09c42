"""Potential mixer on the device: multisecant Type-I, group size 1 (Broyden's
first method), mix 0.5, restart factor 0.1.

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
                 tol: float = float(np.finfo(float).eps)):
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
