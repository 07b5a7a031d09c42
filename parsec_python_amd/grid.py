"""Real-space grid, domain masks and the device index map.

The reference works on the full nx^3 cube with zero Dirichlet boundaries
(``src/Laplacian/fd3d.py:34-106``); grid point ``i`` sits at ``i*h - rad``
(``src/V_ion/pseudoDiag.py:203-212``) and the flat index is ``i + nx*j +
nx*ny*k`` (``src/rsdft_output.py:68``, Fortran order).  Here the domain is any
set of whole x-chords (one contiguous run of points per (j,k) row):

* ``cube``   -- every point (the reference's domain);
* ``sphere`` -- points with x^2+y^2+z^2 <= rad^2 (rad = cube half-width);
* an index-space sphere for synthetic workloads.

Points are stored compactly in flat order, so a vector is an N_domain array and
a block of k vectors is an N x k column-major matrix.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from fractions import Fraction
from math import factorial

import numpy as np

from . import _lib


def fd_weights(radius: int) -> list[Fraction]:
    """Exact weights of -d^2/dx^2 for the central difference of order 2*radius.

    c_k = 2 (-1)^k (p!)^2 / (k^2 (p-k)! (p+k)!), c_0 = -2 sum c_k.  For radius
    1..4 these are the reference's 1D coefficients (fd3d.py:50-94).
    """
    p = radius
    w = [Fraction(0)] * (p + 1)
    for k in range(1, p + 1):
        w[k] = Fraction(2 * (-1) ** k * factorial(p) ** 2, k * k * factorial(p - k) * factorial(p + k))
    w[0] = -2 * sum(w[1:])
    return w


def reference_fd_radius(fd_order: int) -> int:
    """Radius the reference uses for ``fd_order``: orders above 6 give 8th order (fd3d.py:49-79)."""
    if fd_order <= 2:
        return 1
    if fd_order <= 4:
        return 2
    if fd_order <= 6:
        return 3
    return 4


def row_extents_from_mask(mask3: np.ndarray):
    """mask3[k, j, i] -> (lo, hi) per row (flattened j + ny*k); rows must be contiguous."""
    nz, ny, nx = mask3.shape
    m = mask3.reshape(nz * ny, nx)
    any_ = m.any(axis=1)
    lo = np.where(any_, np.argmax(m, axis=1), 0).astype(np.int32)
    hi = np.where(any_, nx - np.argmax(m[:, ::-1], axis=1), 0).astype(np.int32)
    cnt = m.sum(axis=1)
    if not np.array_equal(cnt, (hi - lo)):
        raise ValueError("domain rows must be contiguous chords")
    return lo, hi


@dataclass
class Grid:
    """A domain on an nx*ny*nz grid with spacing h (bohr) and half-width rad."""

    nx: int
    ny: int
    nz: int
    h: float
    rad: float
    mask: np.ndarray            # flat bool over the cube, x fastest
    kind: str

    @property
    def npts(self) -> int:
        return int(self.mask.sum())

    @property
    def flat_index(self) -> np.ndarray:
        return np.nonzero(self.mask)[0]

    def axis(self) -> np.ndarray:
        return np.arange(self.nx) * self.h - self.rad

    def coords(self) -> np.ndarray:
        """(N, 3) coordinates of domain points in storage order."""
        flat = self.flat_index
        x = self.axis()
        i = flat % self.nx
        j = (flat // self.nx) % self.ny
        k = flat // (self.nx * self.ny)
        return np.stack([x[i], x[j], x[k]], axis=1)

    def rows(self):
        return row_extents_from_mask(self.mask.reshape(self.nz, self.ny, self.nx))

    def scatter_to_cube(self, v: np.ndarray) -> np.ndarray:
        out = np.zeros(self.nx * self.ny * self.nz, dtype=np.asarray(v).dtype)
        out[self.mask] = v
        return out

    def gather_from_cube(self, full: np.ndarray) -> np.ndarray:
        return np.asarray(full).reshape(-1)[self.mask]


def make_grid(nx: int, h: float, rad: float, kind: str = "cube") -> Grid:
    n = nx * nx * nx
    if kind == "cube":
        mask = np.ones(n, dtype=bool)
    elif kind == "sphere":
        x = np.arange(nx) * h - rad
        d2 = x[None, None, :] ** 2 + x[None, :, None] ** 2 + x[:, None, None] ** 2
        mask = (d2 <= rad * rad).reshape(-1)
    else:
        raise ValueError(f"unknown domain kind {kind!r}")
    return Grid(nx=nx, ny=nx, nz=nx, h=h, rad=rad, mask=mask, kind=kind)


def make_index_sphere_grid(n: int, center: float, radius_pts: float, h: float) -> Grid:
    """n^3 cube, points with (i-c)^2+(j-c)^2+(k-c)^2 <= r^2 (synthetic workload)."""
    i = np.arange(n) - center
    d2 = i[None, None, :] ** 2 + i[None, :, None] ** 2 + i[:, None, None] ** 2
    mask = (d2 <= radius_pts * radius_pts).reshape(-1)
    return Grid(nx=n, ny=n, nz=n, h=h, rad=center * h, mask=mask, kind="index_sphere")


class DeviceGrid:
    """Owns the C-ABI grid handle (row table + stencil weights) for one radius."""

    def __init__(self, grid: Grid, radius: int):
        lo, hi = grid.rows()
        w = np.array([float(c) for c in fd_weights(radius)] + [0.0] * (8 - radius - 1), dtype=np.float64)
        handle = C.c_void_p()
        npts = C.c_longlong()
        lo = np.ascontiguousarray(lo, dtype=np.int32)
        hi = np.ascontiguousarray(hi, dtype=np.int32)
        _lib.call("pc_grid_create", grid.nx, grid.ny, grid.nz,
                  lo.ctypes.data_as(C.POINTER(C.c_int)), hi.ctypes.data_as(C.POINTER(C.c_int)), radius,
                  w.ctypes.data_as(C.POINTER(C.c_double)), C.byref(handle), C.byref(npts))
        if npts.value != grid.npts:
            raise _lib.ParsecHipError("device grid point count mismatch")
        self.grid = grid
        self.radius = radius
        self.handle = handle
        self.npts = npts.value

    def close(self):
        if self.handle:
            _lib.call("pc_grid_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
