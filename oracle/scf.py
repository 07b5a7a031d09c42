"""SCF driver (oracle restatement of ``src/rsdft_solver.py:156-462``).

Units follow the reference (P4): H = 0.5*A + V_nl + diag(pot) in Hartree with
``A = fd3d/h^2`` (rsdft_solver.py:186,263,285), pot = V_loc + V_H0 + V_H +
0.5*v_xc (v_xc in Ry, :257,351), energies in Ry (:128-153).  Iteration 1 calls
``first_filt(nev+15, ...)``, later iterations ``chefsi1`` (diagmeth 3,
:303-316); density ``rho = sum_i 2 f_i W_i^2`` (:324); Hartree CG warm-started
from the previous V_H with tol 1e-5 and 200 iterations (:327-345); the
reference stops when ``|V_new - V|/|V_new| <= tol`` (:281,352).  Optionally the
run also stops on ``|E_i - E_{i-1}| < etol`` Ha/atom (the metric's criterion).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from . import eig, fd, grid, pseudo, xc
from .mixer import Msecant1
from .pcg import pcg


@dataclass
class Setup:
    domain: dict
    mask: np.ndarray
    A: sp.csr_matrix          # -lap (1/h^2) on masked points
    vnl: sp.csr_matrix
    proj: list
    rho0: np.ndarray
    hpot0: np.ndarray
    ppot: np.ndarray
    e_nuc: float
    nelec: float
    n_atoms: int
    diag_info: dict
    fd_radius: int


def build_setup(atoms, domain, mask_kind="cube", fd_order=8, fd_radius=None) -> Setup:
    mask = grid.mask_flat(domain, mask_kind)
    r = fd.reference_radius(fd_order) if fd_radius is None else fd_radius
    h = domain["h"]
    A = (1.0 / (h * h)) * fd.laplacian_masked(domain["nx"], domain["ny"], domain["nz"], r, mask)
    rho0, hpot0, ppot, info = pseudo.pseudo_diag(domain, atoms, mask)
    proj = pseudo.projector_tables(domain, atoms, mask)
    vnl = pseudo.pseudo_nl(domain, atoms, mask)
    return Setup(domain=domain, mask=mask, A=A.tocsr(), vnl=vnl, proj=proj, rho0=rho0, hpot0=hpot0,
                 ppot=ppot, e_nuc=pseudo.nuclear(atoms), nelec=pseudo.nelectrons(atoms),
                 n_atoms=sum(c.shape[0] for _, c in atoms), diag_info=info, fd_radius=r)


def energy_components(rho, hpot, hpot0, vxc, exc, lam, occ, nev, e_nuc, h):
    """rsdft_solver.py:128-153 (rho in e/bohr^3; result in Ry)."""
    esum = float(np.sum(lam[:nev] * occ[:nev]))
    eig_ry = 4.0 * esum
    hart_ry = float(np.sum(rho * (hpot + hpot0)) * h ** 3)
    vxc_ry = float(np.sum(rho * vxc) * h ** 3)
    total = eig_ry - hart_ry + float(exc) - vxc_ry + float(e_nuc)
    return {"eigen_sum_ry": eig_ry, "hartree_ry": hart_ry, "xc_ry": float(exc), "ion_ry": float(e_nuc),
            "total_ry": total}


@dataclass
class SCFResult:
    energies: dict
    eigenvalues: np.ndarray
    occupations: np.ndarray
    rho: np.ndarray            # e/bohr^3
    hpot: np.ndarray
    pot: np.ndarray
    W: np.ndarray
    iterations: int
    error: float
    converged: bool
    history: list = field(default_factory=list)
    timings: dict = field(default_factory=dict)


def run_scf(setup: Setup, nev: int, *, tol=2e-4, maxits=40, poldeg=10, fermi_temp=500.0,
            seed=0, etol_ha_per_atom=None, max_wall=None, pcg_aliasing=True,
            rho_init=None, hpot_init=None, rng=None, xc_loop=False) -> SCFResult:
    """Run the reference SCF loop.  ``rho_init``/``hpot_init`` replace (rho0, V_H0)
    for a learned/ML start (pseudoDiag_ML4Den_poisson.py contract).  ``rng``
    supplies ``random``/``standard_normal`` start vectors (default: PCG64(seed)).
    ``xc_loop`` evaluates the XC point by point as the reference does (same
    values; used where the reference's cost is timed)."""
    exc_fn = xc.exc_nspn_loop if xc_loop else xc.exc_nspn
    if rng is None:
        rng = np.random.default_rng(seed)
    h = setup.domain["h"]
    n = setup.A.shape[0]
    rho0 = setup.rho0 if rho_init is None else rho_init
    hpot0 = setup.hpot0 if hpot_init is None else hpot_init
    hpot = np.zeros(n)
    vxc, exc = exc_fn(h, rho0 / h ** 3)
    pot = setup.ppot + hpot0 + 0.5 * vxc
    mixer = Msecant1()
    half_a_plus_vnl = (0.5 * setup.A + setup.vnl).tocsr()
    W = lam = occ = None
    err = 10.0 + tol
    its = 0
    hist = []
    e_prev = None
    t_start = time.perf_counter()
    tim = {"diag": 0.0, "hartree": 0.0, "other": 0.0}
    energies = None
    converged = False
    rho = np.zeros(n)
    while err > tol and its < maxits:
        its += 1
        B = (half_a_plus_vnl + sp.diags(pot, 0, shape=(n, n))).tocsr()
        t0 = time.perf_counter()
        if its == 1:
            W, lam, _ = eig.first_filt(nev + 15, B, poldeg, rng)
        else:
            W, lam = eig.chefsi1(W, lam, poldeg, nev, B, rng)
        t1 = time.perf_counter()
        _, occ = eig.occupations(lam[:nev], fermi_temp, setup.nelec, 1e-6)
        rho = (W[:, :nev] ** 2) @ (2 * occ)
        hrhs = (4 * np.pi / h ** 3) * (rho - rho0)
        rho = rho / h ** 3
        hpot, cg_its = pcg(setup.A, hrhs, hpot, 200, 1e-5, reference_aliasing=pcg_aliasing)
        t2 = time.perf_counter()
        vxc, exc = exc_fn(h, rho)
        pot_new = setup.ppot + 0.5 * vxc + hpot + hpot0
        err = float(np.linalg.norm(pot_new - pot) / np.linalg.norm(pot_new))
        energies = energy_components(rho, hpot, hpot0, vxc, exc, lam, occ, nev, setup.e_nuc, h)
        e_ha_atom = 0.5 * energies["total_ry"] / setup.n_atoms
        de = None if e_prev is None else abs(e_ha_atom - e_prev)
        hist.append({"iter": its, "err": err, "total_ry": energies["total_ry"], "dE_ha_atom": de,
                     "cg_its": cg_its, "lam": lam[:nev].copy()})
        e_prev = e_ha_atom
        pot, _ = mixer(pot, pot_new - pot)
        t3 = time.perf_counter()
        tim["diag"] += t1 - t0
        tim["hartree"] += t2 - t1
        tim["other"] += t3 - t2
        if etol_ha_per_atom is not None and de is not None and de < etol_ha_per_atom:
            converged = True
            break
        if max_wall is not None and time.perf_counter() - t_start > max_wall:
            break
    if not converged:
        converged = err <= tol
    tim["total"] = time.perf_counter() - t_start
    return SCFResult(energies=energies, eigenvalues=np.asarray(lam), occupations=np.asarray(occ), rho=rho,
                     hpot=hpot, pot=pot, W=W, iterations=its, error=err, converged=converged, history=hist,
                     timings=tim)
