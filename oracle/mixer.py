"""Density/potential mixer: multisecant Type-I with group size 1 (Broyden's first
method), mix 0.5, restart factor 0.1, EN_like 0 (oracle restatement of
``src/Mixer/msecant1.py:40-175`` with parameters ``src/Mixer/includemix.py:21-49``
and the quasi-inverse ``src/Mixer/qinv.py:5-62``; dispatch ``src/Mixer/mixer.py``).

With group size 1 and EN_like 0 every call after the first:
  * restarts from the last stored iterate when ``|f_prev| < 0.1 |f_new|`` (m >= 2);
  * turns the last stored (x, f) into (dx, df), updates
    ``DX[:,m-1] = dx + mix df - sum_{i<m-1} DX_i (DF_i . df)``;
  * ``N = -mix dx + sum_{i<m-1} DF_i (DX_i . dx)``; ``DF[:,m-1] = N / (N . df)``;
  * ``x_new = x + mix f - sum_{i<m} DX_i (DF_i . f)``; appends (x, f).
"""
from __future__ import annotations

import numpy as np


class Msecant1:
    def __init__(self, mix: float = 0.5, restart_factor: float = 0.1, tol: float = np.finfo(float).eps):
        self.mix = mix
        self.restart_factor = restart_factor
        self.tol = tol
        self.DX = None
        self.DF = None

    def __call__(self, x1, f1):
        mix = self.mix
        x1 = np.asarray(x1, dtype=np.float64).reshape(-1)
        f1 = np.asarray(f1, dtype=np.float64).reshape(-1)
        m = 0 if self.DX is None else len(self.DX)
        if m == 0:
            self.DX = [x1.copy()]
            self.DF = [f1.copy()]
            return x1 + mix * f1, 0
        if m >= 2:
            if np.linalg.norm(self.DF[m - 1]) < self.restart_factor * np.linalg.norm(f1):
                xr, fr = self.DX[m - 1].copy(), self.DF[m - 1].copy()
                self.DX, self.DF = [xr], [fr]
                return xr + mix * fr, 0
        ngroup = m - 1
        dxl = x1 - self.DX[m - 1]
        dfl = f1 - self.DF[m - 1]
        dx = dxl.copy()
        v = dxl + mix * dfl
        for i in range(ngroup):
            v = v - self.DX[i] * float(self.DF[i] @ dfl)
        self.DX[m - 1] = v
        self.DF[m - 1] = dfl
        x_new = x1 + mix * f1
        for i in range(ngroup):
            x_new = x_new - self.DX[i] * float(self.DF[i] @ f1)
        nvec = -mix * dx
        for i in range(ngroup):
            nvec = nvec + self.DF[i] * float(self.DX[i] @ dx)
        mval = float(nvec @ dfl)
        c = _qinv_scalar(mval, self.tol)
        self.DF[m - 1] = nvec * c
        x_new = x_new - self.DX[m - 1] * float(self.DF[m - 1] @ f1)
        self.DX.append(x1.copy())
        self.DF.append(f1.copy())
        return x_new, m


def _qinv_scalar(a: float, tol: float) -> float:
    """qinv of a 1x1 matrix (qinv.py:29-62): 1/a unless |a| < tol*|a| (never for a != 0)."""
    if abs(a) < tol * abs(a):
        return 0.0
    return 1.0 / a
