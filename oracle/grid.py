"""Domain sizing, recentring and masks (oracle restatement).

Follows ``src/rsdft_input.py:582-632`` (unit conversion: ``factor = 1.0/A0_ANG``
then ``coord*factor``), ``src/rsdft_setup.py:19-83`` (h, nev),
``src/rsdft_setup.py:86-150`` (nx, radius) and ``src/rsdft_setup.py:153-160``
(recentring).  Grid point ``i`` sits at ``i*h - rad`` and the flat index is
``i + nx*j + nx*ny*k`` (``src/V_ion/pseudoDiag.py:203-212,273-278``).
"""
from __future__ import annotations

import numpy as np

from .tables import elements

A0_ANG = 0.529177210903  # src/rsdft_models.py:14


def to_bohr(atoms_ang):
    """[(symbol, (n,3) Angstrom)] -> [(symbol, (n,3) bohr)]; rsdft_input.py:611,625."""
    factor = 1.0 / A0_ANG
    return [(s, np.asarray(c, dtype=float).copy() * factor) for s, c in atoms_ang]


def recenter(atoms):
    """Subtract the centroid of all atoms (rsdft_setup.py:153-160)."""
    allc = np.vstack([c for _, c in atoms])
    center = np.mean(allc, axis=0)
    return [(s, c - center) for s, c in atoms]


def size_domain(atoms, h=None, radius=None, nev=None):
    """Return dict(radius, nx, ny, nz, h, nev, zelec) as the reference does.

    ``h``/``radius`` are in bohr (already converted).  ``rsdft_setup.py:45-83``
    picks h = min table h when not given; nev = max(16, round(0.7*Z + 0.5)).
    ``rsdft_setup.py:119-150`` sizes the cube.
    """
    tab = elements()
    zelec = 0.0
    hmin = 100.0
    for s, c in atoms:
        zelec += tab[s]["Z"] * c.shape[0]
        hmin = min(hmin, tab[s]["h"])
    hh = float(h) if h is not None else hmin
    if nev is None:
        nev = max(16, round(0.7 * zelec + 0.5))
    rmax = 0.0
    for s, c in atoms:
        for a in range(c.shape[0]):
            rdis = np.sqrt(c[a, 0] ** 2 + c[a, 1] ** 2 + c[a, 2] ** 2)
            rmax = max(rmax, rdis + tab[s]["r"])
    sph = rmax
    if radius is not None:
        sph = float(radius)
    nx = int(2 * sph / hh) + 1
    nx = 2 * ((nx + 1) // 2)
    rad = 0.5 * hh * (nx - 1)
    return {"radius": rad, "nx": nx, "ny": nx, "nz": nx, "h": hh, "nev": int(nev), "zelec": zelec}


def axis(domain):
    nx, h, rad = domain["nx"], domain["h"], domain["radius"]
    return np.arange(nx) * h - rad


def mask_flat(domain, kind: str):
    """Boolean mask over the flat cube index (x fastest).

    ``cube``: every point (the reference's Hamiltonian size, P1).
    ``sphere``: points with x^2+y^2+z^2 <= rad^2 where rad is the cube
    half-width (the build's performance domain; Dirichlet zero outside).
    """
    nx, ny, nz = domain["nx"], domain["ny"], domain["nz"]
    if kind == "cube":
        return np.ones(nx * ny * nz, dtype=bool)
    if kind != "sphere":
        raise ValueError(f"unknown mask {kind!r}")
    x = axis(domain)
    rad = domain["radius"]
    d2 = x[None, None, :] ** 2 + x[None, :, None] ** 2 + x[:, None, None] ** 2  # [k, j, i]
    return (d2 <= rad * rad).reshape(-1)


def index_sphere_mask(n: int, center: float, radius_pts: float):
    """Mask of an n^3 cube: (i-c)^2+(j-c)^2+(k-c)^2 <= r^2 in index units."""
    i = np.arange(n) - center
    d2 = i[None, None, :] ** 2 + i[None, :, None] ** 2 + i[:, None, None] ** 2
    return (d2 <= radius_pts * radius_pts).reshape(-1)


def point_coords(domain, mask):
    """(N,3) coordinates of masked points in flat order."""
    nx, ny, nz = domain["nx"], domain["ny"], domain["nz"]
    x = axis(domain)
    flat = np.nonzero(mask)[0]
    i = flat % nx
    j = (flat // nx) % ny
    k = flat // (nx * ny)
    return np.stack([x[i], x[j], x[k]], axis=1), flat
