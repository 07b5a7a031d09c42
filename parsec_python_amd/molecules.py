"""Geometries (Angstrom) and the reference's domain sizing.

``size_domain`` follows ``src/rsdft_input.py:582-632`` (Angstrom -> bohr with
``factor = 1.0/A0_ANG`` then ``coord*factor``), ``src/rsdft_setup.py:19-83``
(default h, nev = max(16, round(0.7 Z + 0.5))), ``src/rsdft_setup.py:86-150``
(nx = int(2 r/h) + 1 rounded up to even, rad = h (nx-1)/2) and
``src/rsdft_setup.py:153-160`` (recentring on the centroid).
"""
from __future__ import annotations

import numpy as np

from . import tables

A0_ANG = tables.A0_ANG


def to_bohr(atoms_ang):
    factor = 1.0 / A0_ANG
    return [(s, np.asarray(c, dtype=float).copy() * factor) for s, c in atoms_ang]


def recenter(atoms):
    allc = np.vstack([c for _, c in atoms])
    center = np.mean(allc, axis=0)
    return [(s, c - center) for s, c in atoms]


def size_domain(atoms_bohr, h=None, radius=None, nev=None):
    """Return (domain dict {radius, nx, ny, nz, h}, nev, zelec); h/radius in bohr."""
    tab = tables.elements()
    zelec = 0.0
    hmin = 100.0
    for s, c in atoms_bohr:
        zelec += tab[s]["Z"] * c.shape[0]
        hmin = min(hmin, tab[s]["h"])
    hh = float(h) if h is not None else hmin
    if nev is None:
        nev = max(16, round(0.7 * zelec + 0.5))
    rmax = 0.0
    for s, c in atoms_bohr:
        for a in range(c.shape[0]):
            rmax = max(rmax, float(np.sqrt(c[a, 0] ** 2 + c[a, 1] ** 2 + c[a, 2] ** 2)) + tab[s]["r"])
    sph = float(radius) if radius is not None else rmax
    nx = int(2 * sph / hh) + 1
    nx = 2 * ((nx + 1) // 2)
    rad = 0.5 * hh * (nx - 1)
    return {"radius": rad, "nx": nx, "ny": nx, "nz": nx, "h": hh}, int(nev), zelec


def prepare(atoms_ang, h_ang=None, radius_ang=None, nev=None, recenter_atoms=True):
    """Angstrom input -> (atoms_bohr, domain, nev), as the reference's setup does."""
    factor = 1.0 / A0_ANG
    atoms = to_bohr(atoms_ang)
    h = None if h_ang is None else float(h_ang) * factor
    r = None if radius_ang is None else float(radius_ang) * factor
    if recenter_atoms:
        atoms = recenter(atoms)
    domain, nev, _ = size_domain(atoms, h, r, nev)
    return atoms, domain, nev


# ---------------------------------------------------------------- geometries

def h2o_sample():
    """samples/h2o/H2O_sad_diagmeth3_5A_0p2A.in:2-4 (O-H 0.97 A)."""
    return [("O", [[-3.56626, 1.77639, 0.00000]]),
            ("H", [[-2.59626, 1.77639, 0.00000], [-3.88959, 1.36040, -0.81444]])]


def h2o():
    """O at the origin, H at (+-0.757, 0.587, 0) A."""
    return [("O", [[0.0, 0.0, 0.0]]), ("H", [[0.757, 0.587, 0.0], [-0.757, 0.587, 0.0]])]


def h_atom():
    return [("H", [[0.0, 0.0, 0.0]])]


def naphthalene_sample():
    """samples/h8c10/cpu/H8C10_sad_diagmeth3_9A_0p2A.in:2-19."""
    h = [[9.000000, 11.481313, 10.220927], [9.000000, 10.240247, 12.355972], [9.000000, 7.759753, 12.355972],
         [9.000000, 6.518687, 10.220927], [9.000000, 11.481313, 7.779073], [9.000000, 6.518687, 7.779073],
         [9.000000, 7.759753, 5.644028], [9.000000, 10.240247, 5.644028]]
    c = [[9.000000, 10.387757, 10.228803], [9.000000, 9.700099, 11.407084], [9.000000, 8.299901, 11.407084],
         [9.000000, 7.612243, 10.228803], [9.000000, 8.289915, 9.000000], [9.000000, 9.710085, 9.000000],
         [9.000000, 10.387757, 7.771197], [9.000000, 7.612243, 7.771197], [9.000000, 8.299901, 6.592916],
         [9.000000, 9.700099, 6.592916]]
    return [("H", h), ("C", c)]


def naphthalene(cc=1.40, ch=1.09):
    """Planar C10H8 from two regular hexagons (C-C cc, C-H ch) in the xy plane."""
    r = cc  # hexagon circumradius = side
    ang = np.deg2rad(np.arange(6) * 60.0 + 30.0)
    ring = np.stack([r * np.cos(ang), r * np.sin(ang), np.zeros(6)], axis=1)
    shift = np.array([r * np.cos(np.deg2rad(30.0)), 0.0, 0.0])
    left = ring - shift
    right = ring + shift
    carbons = np.vstack([left, right])
    uniq = []
    for p in carbons:
        if not any(np.linalg.norm(p - q) < 1e-6 for q in uniq):
            uniq.append(p)
    carbons = np.array(uniq)
    hyd = []
    for p in carbons:
        nb = [q for q in carbons if 0 < np.linalg.norm(p - q) < cc * 1.1]
        if len(nb) == 2:
            d = 2 * p - nb[0] - nb[1]
            d = d / np.linalg.norm(d)
            hyd.append(p + ch * d)
    return [("H", np.array(hyd)), ("C", carbons)]


def silicon_cluster(a=5.43, sih=1.48, n_si=87):
    """Atom-centred H-passivated diamond Si cluster (Si87H76 for n_si=87)."""
    base = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
    pts = []
    rng = range(-3, 4)
    for i in rng:
        for j in rng:
            for k in rng:
                for b in base:
                    for off in (np.zeros(3), np.array([0.25, 0.25, 0.25])):
                        pts.append((np.array([i, j, k]) + b + off) * a)
    pts = np.unique(np.round(np.array(pts), 6), axis=0)
    d = np.linalg.norm(pts, axis=1)
    order = np.argsort(d, kind="stable")
    shells = np.unique(np.round(d[order], 4))
    chosen = None
    for rcut in shells:
        sel = pts[d <= rcut + 1e-3]
        if len(sel) >= n_si:
            chosen = sel
            break
    if chosen is None or len(chosen) != n_si:
        raise ValueError(f"no atom-centred shell closes at {n_si} Si atoms")
    bond = a * np.sqrt(3) / 4
    hyd = []
    for p in chosen:
        # the four tetrahedral neighbour directions of this sublattice site
        frac = np.round(p / a * 4) % 2  # sublattice parity via (x+y+z) pattern
        sign = 1.0 if int(round(np.sum(np.round(p / a * 4)))) % 4 == 0 else -1.0
        dirs = sign * np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) / np.sqrt(3)
        del frac
        for dv in dirs:
            q = p + bond * dv
            if not np.any(np.linalg.norm(chosen - q, axis=1) < 1e-3):
                hyd.append(p + sih * dv)
    return [("Si", chosen), ("H", np.array(hyd))]


def alkane(n_c: int, rng: np.random.Generator | None = None, bond_jitter=0.05, angle_jitter_deg=10.0,
           cc=1.54, ch=1.09):
    """Zig-zag CnH2n+2 with optional random bond (+-5 %) and angle (+-10 deg) jitter."""
    tet = np.deg2rad(109.47)

    def jb(x):
        return x * (1.0 + (rng.uniform(-bond_jitter, bond_jitter) if rng is not None else 0.0))

    def ja(x):
        return x + (np.deg2rad(rng.uniform(-angle_jitter_deg, angle_jitter_deg)) if rng is not None else 0.0)

    carbons = [np.zeros(3)]
    direction = np.array([np.sin(tet / 2), np.cos(tet / 2), 0.0])
    for i in range(1, n_c):
        half = ja(tet) / 2
        sgn = 1 if i % 2 else -1
        direction = np.array([np.sin(half), sgn * np.cos(half), 0.0])
        carbons.append(carbons[-1] + jb(cc) * direction)
    carbons = np.array(carbons)
    hyd = []
    for i, p in enumerate(carbons):
        nbrs = [carbons[j] for j in (i - 1, i + 1) if 0 <= j < n_c]
        if not nbrs:
            dirs = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) / np.sqrt(3)
        else:
            u = sum((q - p) / np.linalg.norm(q - p) for q in nbrs)
            u = -u / (np.linalg.norm(u) + 1e-12)
            if len(nbrs) == 1:
                b = (nbrs[0] - p) / np.linalg.norm(nbrs[0] - p)
                perp = np.cross(b, [0, 0, 1.0])
                if np.linalg.norm(perp) < 1e-6:
                    perp = np.cross(b, [0, 1.0, 0])
                perp /= np.linalg.norm(perp)
                perp2 = np.cross(b, perp)
                dirs = []
                for t in range(3):
                    phi = 2 * np.pi * t / 3
                    dvec = -b * np.cos(np.pi - tet) + np.sin(np.pi - tet) * (np.cos(phi) * perp + np.sin(phi) * perp2)
                    dirs.append(-dvec if np.dot(-dvec, b) < 0 else dvec)
                dirs = [d if np.dot(d, b) < 0 else -d for d in dirs]
            else:
                w = np.cross(nbrs[0] - p, nbrs[1] - p)
                w /= np.linalg.norm(w)
                dirs = [np.cos(ja(tet) / 2) * u + np.sin(ja(tet) / 2) * w,
                        np.cos(ja(tet) / 2) * u - np.sin(ja(tet) / 2) * w]
        for dv in dirs:
            dv = np.asarray(dv) / np.linalg.norm(dv)
            hyd.append(p + jb(ch) * dv)
    return [("C", carbons), ("H", np.array(hyd))]
