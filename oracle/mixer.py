"""Density/potential mixers: msecant1 (the reference's default), msecant2 and
msecant3 (src/Mixer/mixer.py:37-107 selects among them).

Multisecant Type-I with group size 1 (Broyden's first
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


def _qinv_column_scale(df: np.ndarray, tol: float) -> float:
    """qinv of an N x 1 matrix (qinv.py:29-62) is df^T / |df|^2, or 0 when
    |df| < tol * max|df_i| (rank 0); returns the scalar C2 C2^T = 1/|df|^2."""
    r = float(np.linalg.norm(df))
    if r < tol * float(np.max(np.abs(df))) or r == 0.0:
        return 0.0
    return 1.0 / (r * r)


class Msecant2:
    """Multisecant Type-II, group size 1, EN_like 0 (Broyden's second method):
    oracle restatement of ``src/Mixer/msecant2.py:24-92``.  A restart keeps
    (x1, f1) and sets the stage to 2 (msecant2.py:45-52), after which the
    stage alternates 2 -> 1 -> 2 as the reference's does."""

    def __init__(self, mix: float = 0.5, restart_factor: float = 0.1, tol: float = np.finfo(float).eps):
        self.mix, self.restart_factor, self.tol = mix, restart_factor, tol
        self.DX = self.DF = None
        self.stage = 0

    def __call__(self, x1, f1):
        mix = self.mix
        x1 = np.asarray(x1, dtype=np.float64).reshape(-1)
        f1 = np.asarray(f1, dtype=np.float64).reshape(-1)
        if self.DX is None:
            self.DX, self.DF, self.stage = [x1.copy()], [f1.copy()], 0
            return x1 + mix * f1, 0
        m = len(self.DX)
        if self.stage != 1 and m >= 2 and np.linalg.norm(self.DF[-1]) < self.restart_factor * np.linalg.norm(f1):
            self.DX, self.DF, self.stage = [x1.copy()], [f1.copy()], 2
            return x1 + mix * f1, 0
        ng = m - 1
        if self.stage != 1:
            dfl = f1 - self.DF[m - 1]
            v = x1 - self.DX[m - 1] + mix * dfl
            for i in range(ng):
                v = v - self.DX[i] * float(self.DF[i] @ dfl)
            self.DX[m - 1], self.DF[m - 1] = v, dfl
        x_new = x1 + mix * f1
        for i in range(ng):
            x_new = x_new - self.DX[i] * float(self.DF[i] @ f1)
        if self.stage == 1:
            x_new = x_new - self.DX[m - 1] * float(self.DF[m - 1] @ f1)
        else:
            c = _qinv_column_scale(self.DF[m - 1], self.tol)
            self.DF[m - 1] = self.DF[m - 1] * c
            x_new = x_new - self.DX[m - 1] * float(self.DF[m - 1] @ f1)
        if self.stage != 2:
            self.DX.append(x1.copy())
            self.DF.append(f1.copy())
        self.stage = {1: 2, 2: 1}.get(self.stage, self.stage)
        return x_new, m


class Msecant3:
    """Hybrid multisecant, group size 1, EN_like 0, preferred type 1: oracle
    restatement of ``src/Mixer/msecant3.py:45-150``.  Each update is the
    Type-I (Broyden-I) or the Type-II (Broyden-II) rank-one change, chosen by
    comparing |df.df_prev| |N.df| with |df.df| |dx.dx_prev| (Type I when the
    first is larger, Type II when smaller, the preferred type on a tie or at
    the first update after a (re)start).  A restart drops the whole history
    and starts again from (x1, f1)."""

    def __init__(self, mix: float = 0.5, restart_factor: float = 0.1, tol: float = np.finfo(float).eps,
                 preferred_type: int = 1):
        self.mix, self.restart_factor, self.tol, self.preferred = mix, restart_factor, tol, preferred_type
        self.DX = self.DF = None
        self.dx_prev = self.df_prev = None
        self.last_type = None

    def __call__(self, x1, f1):
        mix = self.mix
        x1 = np.asarray(x1, dtype=np.float64).reshape(-1)
        f1 = np.asarray(f1, dtype=np.float64).reshape(-1)
        if self.DX is None:
            self.DX, self.DF = [x1.copy()], [f1.copy()]
            self.dx_prev = np.zeros_like(x1)
            self.df_prev = np.zeros_like(x1)
            return x1 + mix * f1, 0
        m = len(self.DX)
        if m >= 2 and np.linalg.norm(self.DF[-1]) < self.restart_factor * np.linalg.norm(f1):
            self.DX = None
            return self(x1, f1)
        ng = m - 1
        dfl = f1 - self.DF[m - 1]
        dx = x1 - self.DX[m - 1]
        v = dx + mix * dfl
        for i in range(ng):
            v = v - self.DX[i] * float(self.DF[i] @ dfl)
        self.DX[m - 1], self.DF[m - 1] = v, dfl
        x_new = x1 + mix * f1
        for i in range(ng):
            x_new = x_new - self.DX[i] * float(self.DF[i] @ f1)
        nvec = -mix * dx
        for i in range(ng):
            nvec = nvec + self.DF[i] * float(self.DX[i] @ dx)
        if ng > 0:
            lhs = abs(float(dfl @ self.df_prev)) * abs(float(nvec @ dfl))
            rhs = abs(float(dfl @ dfl)) * abs(float(dx @ self.dx_prev))
            typ = 1 if lhs > rhs else (2 if lhs < rhs else self.preferred)
        else:
            typ = self.preferred
        self.dx_prev, self.df_prev = dx.copy(), dfl.copy()
        if typ == 1:
            c = _qinv_scalar(float(nvec @ dfl), self.tol)
            self.DF[m - 1] = nvec * c
        else:
            self.DF[m - 1] = dfl * _qinv_column_scale(dfl, self.tol)
        x_new = x_new - self.DX[m - 1] * float(self.DF[m - 1] @ f1)
        self.DX.append(x1.copy())
        self.DF.append(f1.copy())
        self.last_type = typ
        return x_new, m
