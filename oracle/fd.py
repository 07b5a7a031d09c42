"""Finite-difference Laplacian (oracle restatement of ``src/Laplacian/fd3d.py``).

The reference builds ``-lap * h^2`` from hard-coded 1D coefficients for orders
2/4/6/8 (``fd3d.py:49-94``) as a Kronecker sum truncated at the cube faces
(zero Dirichlet: the diagonals ``X1..X4``/``Y1..Y4`` carry zeros at row ends),
and any ``degree > 6`` gives 8th order (P3).  Orders above 8 (the synthetic
workload's order 12) come from the closed-form central-difference weights
(Fornberg); the reference does not implement them.
"""
from __future__ import annotations

from fractions import Fraction
from math import factorial

import numpy as np
import scipy.sparse as sp

# The reference's own 1D weights of -d2/dx2 (fd3d.py:50-94): centre, then
# offsets 1..p.  Used only to pin the closed form.
REFERENCE_WEIGHTS = {
    1: [Fraction(2), Fraction(-1)],
    2: [Fraction(5, 2), Fraction(-4, 3), Fraction(1, 12)],
    3: [Fraction(49, 18), Fraction(-3, 2), Fraction(3, 20), Fraction(-1, 90)],
    4: [Fraction(205, 72), Fraction(-8, 5), Fraction(1, 5), Fraction(-8, 315), Fraction(1, 560)],
}


def reference_radius(degree: int) -> int:
    """Stencil radius the reference uses for ``fd_order`` (fd3d.py:49,55,65,79)."""
    if degree <= 2:
        return 1
    if degree <= 4:
        return 2
    if degree <= 6:
        return 3
    return 4


def fd_coefficients(radius: int):
    """Exact weights of -d2/dx2 at offsets 0..radius (order 2*radius).

    c_k = 2 (-1)^k (p!)^2 / (k^2 (p-k)! (p+k)!) for k >= 1 (weights of +d2/dx2
    negated), c_0 = -2 sum_k c_k.
    """
    p = radius
    w = [Fraction(0)] * (p + 1)
    for k in range(1, p + 1):
        w[k] = Fraction(2 * (-1) ** k * factorial(p) ** 2, k * k * factorial(p - k) * factorial(p + k))
    w[0] = -2 * sum(w[1:])
    return w


def laplacian_1d(n: int, radius: int) -> sp.csr_matrix:
    w = [float(c) for c in fd_coefficients(radius)]
    diags = [np.full(n - abs(d), w[abs(d)]) for d in range(-radius, radius + 1) if n - abs(d) > 0]
    offs = [d for d in range(-radius, radius + 1) if n - abs(d) > 0]
    return sp.diags(diags, offs, shape=(n, n), format="csr")


def laplacian(nx: int, ny: int, nz: int, radius: int) -> sp.csr_matrix:
    """-lap*h^2 on the full cube, flat index i + nx*j + nx*ny*k (x fastest)."""
    tx, ty, tz = laplacian_1d(nx, radius), laplacian_1d(ny, radius), laplacian_1d(nz, radius)
    ix, iy, iz = sp.identity(nx, format="csr"), sp.identity(ny, format="csr"), sp.identity(nz, format="csr")
    a = sp.kron(iz, sp.kron(iy, tx)) + sp.kron(iz, sp.kron(ty, ix)) + sp.kron(tz, sp.kron(iy, ix))
    return a.tocsr()


def laplacian_masked(nx, ny, nz, radius, mask) -> sp.csr_matrix:
    """Restriction to masked points: zero Dirichlet outside the mask."""
    a = laplacian(nx, ny, nz, radius)
    if mask.all():
        return a
    idx = np.nonzero(mask)[0]
    return a[idx][:, idx].tocsr()
