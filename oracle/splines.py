"""Radial cubic splines (oracle restatement).

* ``preprocess``: ``src/V_ion/preProcess.py:4-24`` -- keep index 0, every i in
  1..n-2 with ``|v[i]-v[i-1]| > 10e-6``, and the last index.
* ``spline``: ``src/Splines/fspline.py:5-45`` + ``trid.py`` -- natural cubic
  spline; ``z`` is the second derivative divided by 6, and
  ``c = y_j/h_j - z_j h_j``, ``d = y_{j+1}/h_j - z_{j+1} h_j``.
* ``evaluate``: ``src/Splines/fsplevalIO.py:2-65`` -- clamp x into
  [xi_0, xi_last], find the bracketing interval, evaluate
  ``t1 (z_j t1^2/h + c_j) + t2 (z_{j+1} t2^2/h + d_j)``.  The reference's
  interval hint only selects which of two equal-valued adjacent intervals is
  used at a knot, so a bracketing search gives the same values.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import solve_banded

PREPROCESS_EPS = 10e-6


def preprocess(v: np.ndarray) -> np.ndarray:
    v = np.asarray(v, dtype=np.float64).reshape(-1)
    n = len(v)
    inner = np.nonzero(np.abs(v[1:n - 1] - v[0:n - 2]) > PREPROCESS_EPS)[0] + 1
    return np.concatenate([[0], inner, [n - 1]]).astype(np.int64)


def spline(xi: np.ndarray, yi: np.ndarray):
    xi = np.asarray(xi, dtype=np.float64).reshape(-1)
    yi = np.asarray(yi, dtype=np.float64).reshape(-1)
    n = len(xi) - 1
    h = xi[1:] - xi[:-1]
    r = (yi[1:] - yi[:-1]) / h
    r = r[1:] - r[:-1]
    lower = h[:-1]
    upper = h[1:]
    diag = 2.0 * (lower + upper)
    m = n - 1
    ab = np.zeros((3, m))
    ab[0, 1:] = upper[:-1]   # super-diagonal of row k-1 is h[k]
    ab[1, :] = diag
    ab[2, :-1] = lower[1:]   # sub-diagonal of row k is h[k]
    zin = solve_banded((1, 1), ab, r)
    z = np.concatenate([[0.0], zin, [0.0]])
    c = yi[:-1] / h - z[:-1] * h
    d = yi[1:] / h - z[1:] * h
    return z, c, d


def evaluate(xi, z, c, d, x):
    xi = np.asarray(xi)
    x = np.clip(np.asarray(x, dtype=np.float64), xi[0], xi[-1])
    j = np.searchsorted(xi, x, side="right") - 1
    j = np.clip(j, 0, len(xi) - 2)
    t1 = xi[j + 1] - x
    t2 = x - xi[j]
    hj = xi[j + 1] - xi[j]
    return t1 * (z[j] * t1 * t1 / hj + c[j]) + t2 * (z[j + 1] * t2 * t2 / hj + d[j])


class RadialSpline:
    """A preprocessed natural spline of one table column against radius."""

    def __init__(self, radius: np.ndarray, values: np.ndarray):
        keep = preprocess(values)
        self.xi = np.asarray(radius, dtype=np.float64)[keep]
        self.yi = np.asarray(values, dtype=np.float64)[keep]
        self.z, self.c, self.d = spline(self.xi, self.yi)

    def __call__(self, r):
        return evaluate(self.xi, self.z, self.c, self.d, r)
