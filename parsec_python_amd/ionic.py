"""Ionic setup: SAD density, local and atomic-Hartree potentials (device
kernel), Kleinman-Bylander projector tables, ion-ion energy, electron count.

* ``pseudo_diag``       -- ``src/V_ion/pseudoDiag.py:44-421`` (device: pc_sad_eval)
* ``projector_tables``  -- ``src/V_ion/pseudoNL_original.py:9-208`` in factored form
* ``nuclear``           -- ``src/Laplacian/nuclear.py:26-91``
* ``nelectrons``        -- ``src/V_ion/nelectrons.py``

Splines follow ``src/V_ion/preProcess.py`` (knot decimation),
``src/Splines/fspline.py`` (natural cubic, z = y''/6) and
``src/Splines/fsplevalIO.py`` (clamped evaluation).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch
from scipy.linalg import solve_banded

from . import _lib, tables
from .grid import DeviceGrid, Grid

PREPROCESS_EPS = 10e-6


def decimate(v: np.ndarray) -> np.ndarray:
    """Indices kept by the reference's preProcess: 0, every i with |v[i]-v[i-1]| > 1e-5, last."""
    n = len(v)
    keep = np.nonzero(np.abs(v[1:n - 1] - v[0:n - 2]) > PREPROCESS_EPS)[0] + 1
    return np.concatenate([[0], keep, [n - 1]])


def natural_spline(xi: np.ndarray, yi: np.ndarray):
    h = np.diff(xi)
    r = np.diff(np.diff(yi) / h)
    m = len(r)
    ab = np.zeros((3, m))
    ab[0, 1:] = h[1:m]
    ab[1, :] = 2.0 * (h[:-1] + h[1:])
    ab[2, :-1] = h[1:m]
    z = np.concatenate([[0.0], solve_banded((1, 1), ab, r), [0.0]])
    c = yi[:-1] / h - z[:-1] * h
    d = yi[1:] / h - z[1:] * h
    return z, c, d


class Spline:
    def __init__(self, radius: np.ndarray, values: np.ndarray):
        keep = decimate(values)
        self.xi = np.ascontiguousarray(radius[keep], dtype=np.float64)
        yi = np.ascontiguousarray(values[keep], dtype=np.float64)
        self.z, self.c, self.d = natural_spline(self.xi, yi)

    def host_eval(self, r: np.ndarray) -> np.ndarray:
        x = np.clip(r, self.xi[0], self.xi[-1])
        j = np.clip(np.searchsorted(self.xi, x, side="right") - 1, 0, len(self.xi) - 2)
        t1 = self.xi[j + 1] - x
        t2 = x - self.xi[j]
        hj = self.xi[j + 1] - self.xi[j]
        return t1 * (self.z[j] * t1 * t1 / hj + self.c[j]) + t2 * (self.z[j + 1] * t2 * t2 / hj + self.d[j])


def normalize_atoms(atoms):
    """Accept [(sym, coords)] or the reference's [{'typ', 'coord'}] (bohr)."""
    out = []
    for a in atoms:
        if isinstance(a, dict):
            out.append((str(a["typ"]), np.asarray(a["coord"], dtype=np.float64).reshape(-1, 3)))
        else:
            out.append((str(a[0]), np.asarray(a[1], dtype=np.float64).reshape(-1, 3)))
    return out


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: torch.Tensor):
    return C.c_void_p(t.data_ptr())


def pseudo_diag(grid: Grid, dgrid: DeviceGrid, atoms):
    """Return (rho0 [e/cell], V_H0 [Ha], V_loc [Ha], info) as FP64 device tensors."""
    atoms = normalize_atoms(atoms)
    tab = tables.elements()
    species = sorted({s for s, _ in atoms})
    xi_all, z_all, c_all, d_all, off, ln = [], [], [], [], [], []
    cur = 0
    for s in species:
        t = tables.table(s)
        for col in (tables.CHARGE, tables.POT_S, tables.HARTREE):
            sp = Spline(t[:, tables.RADIUS], t[:, col])
            # c, d have one entry per interval (n-1); pad to the knot count so one offset serves all four
            xi_all.append(sp.xi); z_all.append(sp.z)
            c_all.append(np.append(sp.c, 0.0)); d_all.append(np.append(sp.d, 0.0))
            off.append(cur); ln.append(len(sp.xi)); cur += len(sp.xi)
    pos, spec = [], []
    zsum = 0.0
    for s, xyz in atoms:
        zsum += tab[s]["Z"] * xyz.shape[0]
        for a in range(xyz.shape[0]):
            pos.append(xyz[a]); spec.append(species.index(s))
    dev = torch.device("cuda")
    f64 = dict(dtype=torch.float64, device=dev)
    xi_t = torch.tensor(np.concatenate(xi_all), **f64)
    z_t = torch.tensor(np.concatenate(z_all), **f64)
    c_t = torch.tensor(np.concatenate(c_all), **f64)
    d_t = torch.tensor(np.concatenate(d_all), **f64)
    off_t = torch.tensor(off, dtype=torch.int32, device=dev)
    ln_t = torch.tensor(ln, dtype=torch.int32, device=dev)
    pos_t = torch.tensor(np.array(pos, dtype=np.float64).reshape(-1), **f64)
    spec_t = torch.tensor(spec, dtype=torch.int32, device=dev)
    n = grid.npts
    pot = torch.empty(n, **f64)
    rho = torch.empty(n, **f64)
    hpot0 = torch.empty(n, **f64)
    _lib.call("pc_sad_eval", dgrid.handle, grid.h, grid.rad, len(pos), _ptr(pos_t), _ptr(spec_t), _ptr(off_t),
              _ptr(ln_t), _ptr(xi_t), _ptr(z_t), _ptr(c_t), _ptr(d_t), _ptr(pot), _ptr(rho), _ptr(hpot0), _stream())
    s0 = float(rho.sum().item())
    rho0 = rho * (zsum / s0)
    info = {"electron_count_initial": s0, "electron_count_normalized": float(rho0.sum().item()),
            "electron_target": float(zsum)}
    return rho0, hpot0, pot, info


def projector_tables(grid: Grid, atoms):
    """Factored KB p-projectors on the domain (pseudoNL_original.py:60-206).

    Returns (atom_ptr int32[natoms+1], ent_pt int32[nent], ent_u float64[nent,3],
    scale float64[natoms]) with V_nl = sum_a U_a diag(scale_a) U_a^T,
    scale = h^3/Zvalue, u_m = (x_m/d) wfn_P(d) (pot_P-pot_S)(d) for 0 < d < R
    inside the window nearest-index +- round(R/h).
    """
    atoms = normalize_atoms(atoms)
    nx, ny, nz, h, rad = grid.nx, grid.ny, grid.nz, grid.h, grid.rad
    tab = tables.elements()
    remap = -np.ones(nx * ny * nz, dtype=np.int64)
    remap[grid.mask] = np.arange(grid.npts)
    ptr, pts, us, scales = [0], [], [], []
    for s, xyz in atoms:
        t = tables.table(s)
        wfn = Spline(t[:, tables.RADIUS], t[:, tables.WFN_P])
        pps = Spline(t[:, tables.RADIUS], t[:, tables.POT_P] - t[:, tables.POT_S])
        xint = tab[s]["Zvalue"] / h ** 3
        rz = tab[s]["R"]
        span = int(round(rz / h))
        for a in range(xyz.shape[0]):
            xa, ya, za = xyz[a]
            c0 = [int(round((q + rad) / h + 1)) for q in (xa, ya, za)]
            lo = [max(1, c - span) for c in c0]
            hi = [min(nn, c + span) for c, nn in zip(c0, (nx, ny, nz))]
            if any(l > u for l, u in zip(lo, hi)):
                continue
            kk, jj, ii = np.meshgrid(np.arange(lo[2], hi[2] + 1), np.arange(lo[1], hi[1] + 1),
                                     np.arange(lo[0], hi[0] + 1), indexing="ij")
            ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
            dx = (ii - 1) * h - rad - xa
            dy = (jj - 1) * h - rad - ya
            dz = (kk - 1) * h - rad - za
            dd = np.sqrt(dx ** 2 + dy ** 2 + dz ** 2)
            sel = (dd > 0) & (dd < rz)
            flat = ((ii - 1) + nx * ((jj - 1) + (kk - 1) * ny))[sel]
            idx = remap[flat]
            keep = idx >= 0
            dxs, dys, dzs, dds = dx[sel][keep], dy[sel][keep], dz[sel][keep], dd[sel][keep]
            w = wfn.host_eval(dds)
            v = pps.host_eval(dds)
            u = np.stack([dxs / dds * w * v, dys / dds * w * v, dzs / dds * w * v], axis=1)
            pts.append(idx[keep].astype(np.int32))
            us.append(u)
            scales.append(1.0 / xint)
            ptr.append(ptr[-1] + int(keep.sum()))
    if pts:
        ent_pt = np.concatenate(pts)
        ent_u = np.concatenate(us, axis=0)
    else:
        ent_pt = np.zeros(0, dtype=np.int32)
        ent_u = np.zeros((0, 3))
    return (np.array(ptr, dtype=np.int32), ent_pt.astype(np.int32), np.ascontiguousarray(ent_u, dtype=np.float64),
            np.array(scales, dtype=np.float64))


class DeviceProjectors:
    """C-ABI handle for the factored projectors."""

    def __init__(self, grid: Grid, atoms):
        ptr, pt, u, sc = projector_tables(grid, atoms)
        self.natoms = len(sc)
        self.nent = int(ptr[-1])
        self.tables = (ptr, pt, u, sc)
        handle = C.c_void_p()
        _lib.call("pc_proj_create", self.natoms, ptr.ctypes.data_as(C.POINTER(C.c_int)),
                  pt.ctypes.data_as(C.POINTER(C.c_int)), u.ctypes.data_as(C.POINTER(C.c_double)),
                  sc.ctypes.data_as(C.POINTER(C.c_double)), grid.npts, C.byref(handle))
        self.handle = handle

    def close(self):
        if self.handle:
            _lib.call("pc_proj_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def nuclear(atoms) -> float:
    """E_nn = 2 sum_{i<j} Z_i Z_j / R_ij (Ry), valence charges (nuclear.py:178-197)."""
    atoms = normalize_atoms(atoms)
    tab = tables.elements()
    pos = np.vstack([xyz for _, xyz in atoms]) if atoms else np.zeros((0, 3))
    z = np.concatenate([[tab[s]["Z"]] * xyz.shape[0] for s, xyz in atoms]) if atoms else np.zeros(0)
    e = 0.0
    for i in range(len(z)):
        d = pos[i + 1:] - pos[i]
        r = np.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2 + d[:, 2] ** 2)
        e += float(np.sum(z[i] * z[i + 1:] / r))
    return 2.0 * e


def nelectrons(atoms) -> float:
    atoms = normalize_atoms(atoms)
    tab = tables.elements()
    return float(sum(tab[s]["Z"] * xyz.shape[0] for s, xyz in atoms))
