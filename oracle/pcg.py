"""Hartree CG solve (oracle restatement of ``src/Eigensolvers/pcg.py:7-85``).

Stops when ``r.r <= tol^2 * r0.r0`` (relative to the *initial* residual,
pcg.py:49-56) or after ``m`` iterations.  Without a preconditioner the
reference sets ``z = r`` and ``p = z`` -- the same array -- so the in-place
``r -= alp*ap`` of the first iteration also updates ``p``, and the second search
direction becomes ``(1+beta) r_1`` (pcg.py:45-77; the CuPy twin copies and
does not do this).  ``reference_aliasing=True`` reproduces that.
"""
from __future__ import annotations

import numpy as np


def pcg(A, rhs, x0, m, tol, reference_aliasing: bool = True):
    x = np.array(x0, dtype=np.float64, copy=True).reshape(-1)
    r = np.asarray(rhs, dtype=np.float64).reshape(-1) - A @ x
    p = r if reference_aliasing else r.copy()
    ro1 = float(np.dot(r, r))
    tol1 = tol * tol * ro1
    its = 0
    while its < m and ro1 > tol1:
        its += 1
        ro = ro1
        ap = A @ p
        alp = ro / float(np.dot(ap, p))
        x += alp * p
        r -= alp * ap
        ro1 = float(np.dot(r, r))
        bet = ro1 / ro
        p = r + bet * p
    return x, its
