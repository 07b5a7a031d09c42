"""Ionic setup: SAD density, local and atomic-Hartree potentials, KB projectors,
ion-ion energy, electron count (oracle restatement).

* ``pseudo_diag``  follows ``src/V_ion/pseudoDiag.py:44-421``
* ``pseudo_nl``    follows ``src/V_ion/pseudoNL_original.py:9-208``
* ``nuclear``      follows ``src/Laplacian/nuclear.py:26-91``
* ``nelectrons``   follows ``src/V_ion/nelectrons.py``
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import tables
from .grid import axis
from .splines import RadialSpline

T = tables


def _splines_diag(symbol):
    d = T.spline_table(symbol)
    r = d[:, T.COL_RADIUS]
    return (RadialSpline(r, d[:, T.COL_CHARGE]), RadialSpline(r, d[:, T.COL_POT_S]),
            RadialSpline(r, d[:, T.COL_HARTREE]))


def _splines_nl(symbol):
    d = T.spline_table(symbol)
    r = d[:, T.COL_RADIUS]
    pot_ps = d[:, T.COL_POT_P] - d[:, T.COL_POT_S]  # pseudoNL_original.py:76-83
    return RadialSpline(r, d[:, T.COL_WFN_P]), RadialSpline(r, pot_ps)


def pseudo_diag(domain, atoms, mask):
    """Return (rho0 [e/cell], V_H0 [Ha], V_loc [Ha], info) on masked points.

    Per atom: V_loc += pot_S(d), V_H0 += hartree(d), rho0 += max(0, charge(d) h^3/4pi)
    (pseudoDiag.py:199-278); then rho0 *= Z_sum / sum(rho0) (pseudoDiag.py:294-296).
    """
    nx, ny, nz, h, rad = domain["nx"], domain["ny"], domain["nz"], domain["h"], domain["radius"]
    x = axis(domain)
    flat = np.nonzero(mask)[0]
    ii, jj, kk = flat % nx, (flat // nx) % ny, flat // (nx * ny)
    n = flat.size
    pot, rho0, hpot0 = np.zeros(n), np.zeros(n), np.zeros(n)
    tab = T.elements()
    zsum = 0.0
    for sym, xyz in atoms:
        zsum += tab[sym]["Z"] * xyz.shape[0]
        s_chg, s_pot, s_vht = _splines_diag(sym)
        for a in range(xyz.shape[0]):
            dz = (x - xyz[a, 2]) ** 2
            dy = (x - xyz[a, 1]) ** 2
            dx = (x - xyz[a, 0]) ** 2
            r1 = np.sqrt(dz[kk] + dy[jj] + dx[ii])
            pot += s_pot(r1)
            rho0 += np.maximum(0.0, s_chg(r1) * (h ** 3) / (4 * np.pi))
            hpot0 += s_vht(r1)
    s0 = float(np.sum(rho0))
    rho0 = zsum * rho0 / s0
    info = {"electron_count_initial": s0, "electron_count_normalized": float(np.sum(rho0)),
            "electron_target": float(zsum)}
    return rho0, hpot0, pot, info


def projector_tables(domain, atoms, mask):
    """KB p-channel projectors, factored (P2).

    Returns a list of dicts ``{idx (npts,) masked-point index, u (npts,3), scale}``
    with ``V_nl = sum_a U_a diag(scale_a) U_a^T``, ``scale = h^3/Zvalue``.
    Support points: window i0 +- round(R/h) around the nearest grid index
    (pseudoNL_original.py:122-138), then 0 < d < R (:154).
    ``u_m = (x_m/d) * wfn_P(d) * (pot_P - pot_S)(d)`` (:187-189).
    """
    nx, ny, nz, h, rad = domain["nx"], domain["ny"], domain["nz"], domain["h"], domain["radius"]
    tab = T.elements()
    remap = -np.ones(nx * ny * nz, dtype=np.int64)
    remap[np.nonzero(mask)[0]] = np.arange(int(mask.sum()))
    out = []
    for sym, xyz in atoms:
        s_wfn, s_pps = _splines_nl(sym)
        xint = tab[sym]["Zvalue"] / (h ** 3)
        rzero = tab[sym]["R"]
        for a in range(xyz.shape[0]):
            xa, ya, za = xyz[a]
            i0 = int(round((xa + rad) / h + 1))
            j0 = int(round((ya + rad) / h + 1))
            k0 = int(round((za + rad) / h + 1))
            span = int(round(rzero / h))
            i_min, i_max = max(1, i0 - span), min(nx, i0 + span)
            j_min, j_max = max(1, j0 - span), min(ny, j0 + span)
            k_min, k_max = max(1, k0 - span), min(nz, k0 + span)
            if i_min > i_max or j_min > j_max or k_min > k_max:
                continue
            kk, jj, ii = np.meshgrid(np.arange(k_min, k_max + 1), np.arange(j_min, j_max + 1),
                                     np.arange(i_min, i_max + 1), indexing="ij")
            kk, jj, ii = kk.ravel(), jj.ravel(), ii.ravel()
            xxx = (ii - 1) * h - rad - xa
            yyy = (jj - 1) * h - rad - ya
            zzz = (kk - 1) * h - rad - za
            dd = np.sqrt(xxx ** 2 + yyy ** 2 + zzz ** 2)
            sel = (dd > 0) & (dd < rzero)
            if not sel.any():
                continue
            xxx, yyy, zzz, dd = xxx[sel], yyy[sel], zzz[sel], dd[sel]
            flat = (ii[sel] - 1) + nx * ((jj[sel] - 1) + (kk[sel] - 1) * ny)
            w = s_wfn(dd)
            v = s_pps(dd)
            u = np.stack([xxx / dd * w * v, yyy / dd * w * v, zzz / dd * w * v], axis=1)
            idx = remap[flat]
            keep = idx >= 0
            if not keep.any():
                continue
            out.append({"idx": idx[keep], "u": u[keep], "scale": 1.0 / xint, "symbol": sym})
    return out


def pseudo_nl(domain, atoms, mask) -> sp.csr_matrix:
    """V_nl as a sparse N x N matrix (the reference's operator, pseudoNL_original.py:192-206)."""
    n = int(mask.sum())
    tabs = projector_tables(domain, atoms, mask)
    if not tabs:
        return sp.csr_matrix((n, n))
    rows, cols, vals, scales = [], [], [], []
    col = 0
    for t in tabs:
        for m in range(3):
            rows.append(t["idx"])
            cols.append(np.full(t["idx"].size, col))
            vals.append(t["u"][:, m])
            scales.append(t["scale"])
            col += 1
    mm = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, col))
    return (mm @ sp.diags(np.array(scales)) @ mm.T).tocsr()


def apply_nl(tabs, x):
    """V_nl @ x using the factored tables (x: (N,) or (N,k))."""
    y = np.zeros_like(x)
    for t in tabs:
        c = t["u"].T @ x[t["idx"]]
        y[t["idx"]] += t["scale"] * (t["u"] @ c)
    return y


def nuclear(atoms) -> float:
    """E_nn = 2 sum_{i<j} Z_i Z_j / R_ij in Ry (nuclear.py:78-90)."""
    tab = T.elements()
    pos, zz = [], []
    for sym, xyz in atoms:
        for a in range(xyz.shape[0]):
            pos.append(xyz[a])
            zz.append(tab[sym]["Z"])
    e = 0.0
    for i in range(len(pos)):
        for j in range(i + 1, len(pos)):
            d = pos[i] - pos[j]
            e += zz[i] * zz[j] / np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
    return 2.0 * e


def nelectrons(atoms) -> float:
    tab = T.elements()
    return float(sum(tab[s]["Z"] * xyz.shape[0] for s, xyz in atoms))
