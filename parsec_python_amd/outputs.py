"""Run outputs in the reference's formats: the ``.out`` text log, the three
density files and the ``wfn.dat`` binary, plus the reader that starts a run
from a saved density (the reference's ML-density mode).

* ``OutLog`` writes the blocks of ``src/rsdft_output.py:40-295`` (header,
  parameter block, setup timings, electron counts, initial-density line,
  one block per SCF iteration, total-energy summary) with the same line
  layout, so a run can be compared line by line with ``samples/*.out``;
  ``parse_out`` reads such a log back (ours or the reference's).
* ``save_density_variants`` writes ``<base>_grid.npy`` (electrons per cell),
  ``<base>_bohr3.npy`` and ``<base>.npy`` (Angstrom^-3), each an
  (nx, ny, nz) array whose Fortran-order flattening is the flat grid index
  i + nx*j + nx*ny*k (``src/rsdft_output.py:49-83``).
* ``save_wavefunction`` / ``read_wavefunction``: the ``wfn.dat`` layout of
  ``src/rsdft_output.py:298-334`` (uint32 26, radius, h, potential, density,
  states column by column, atom blocks).  Arrays are written on the whole
  nx*ny*nz cube; points outside a sphere domain hold zeros.
* ``density_start_from_file``: the start density of
  ``src/V_ion/pseudoDiag_ML4Den_poisson.py:392-540`` read from an Angstrom^-3
  (nx, ny, nz) ``.npy``: Fortran-order flattening onto the grid, times
  a0^3 h^3, negative values set to zero, renormalised to the electron count,
  and V_H1 = V_H0 + CG(A, 4 pi/h^3 (rho1 - rho0)) from zero to the relative
  tolerance 1e-5.
"""
from __future__ import annotations

import math
import os
import re

import numpy as np

from .tables import A0_ANG

RY_EV = 13.605698066


# ---------------------------------------------------------------------------- .out log
class OutLog:
    """Appends the reference's log blocks to ``path`` (created on open)."""

    def __init__(self, path: str, backend_label: str, density_method: str):
        self.path = path
        d = os.path.dirname(path)
        if d:
            os.makedirs(d, exist_ok=True)
        with open(path, "w", encoding="utf-8") as f:
            f.write(f"Solver backend: {backend_label}\n")
            f.write(f"Density initialization: {density_method}\n")

    def _w(self, text: str) -> None:
        with open(self.path, "a", encoding="utf-8") as f:
            f.write(text)

    def parameters(self, nev, atoms, domain, poldeg, fd_order) -> None:
        """atoms: [(symbol, (n, 3) bohr)]."""
        out = ["\n", f" Number of states: \t{nev}\n\n", "Atom data:\n -------------\n",
               f" Total # of atom types is {len(atoms)}\n"]
        count = 0
        for sym, xyz in atoms:
            xyz = np.asarray(xyz, dtype=float)
            out.append(f" There are {xyz.shape[0]} {sym} atoms\n")
            out.append(" and their coordinates are:\n\n")
            out.append("\t{:<12s}\t{:<12s}\t{:<12s}\n".format("x [a.u.]", "y [a.u.]", "z [a.u.]"))
            count += xyz.shape[0]
            for r in xyz:
                out.append(f"\t{r[0]:12.6f}\t{r[1]:12.6f}\t{r[2]:12.6f}\n")
            out.append("\n")
            out.append("\t{:<12s}\t{:<12s}\t{:<12s}\n".format("x [Ang]", "y [Ang]", "z [Ang]"))
            for r in xyz * A0_ANG:
                out.append(f"\t{r[0]:12.6f}\t{r[1]:12.6f}\t{r[2]:12.6f}\n")
            out.append("\n")
        h = domain["h"]
        out += [" --------------------------------------------------\n",
                f" Total number of atoms :         {count}\n\n",
                f" Number of states:               {nev:10d} \n",
                f" h grid spacing :                {h:10.5f} a.u.   ({h * A0_ANG:10.5f} Ang)\n",
                f' Hamiltonian size :              {domain["nx"] * domain["ny"] * domain["nz"]:10d}  \n',
                f' Sphere Radius :                 {domain["radius"]:10.5f} a.u.   '
                f'({domain["radius"] * A0_ANG:10.5f} Ang)\n',
                f' # grid points in each direction {domain["nx"]:10d}  \n',
                f" Polynomial degree used :        {poldeg:10d}  \n",
                f" Finite difference order :       {fd_order:10d}  \n",
                " --------------------------------------------------\n"]
        self._w("".join(out))

    def electron_counts(self, info: dict) -> None:
        if not info:
            return
        out = [" --------------------------------------------------\n", " Initial density electron count\n"]
        if info.get("electron_count_initial") is not None:
            out.append(f" Initial density integrates to :     {info['electron_count_initial']:.6f} e-\n")
        if info.get("electron_count_normalized") is not None:
            tgt = info.get("electron_target")
            out.append(" Normalized density integrates to:  "
                       + (f"{info['electron_count_normalized']:.6f} e- (target {tgt:.6f})\n" if tgt is not None
                          else f"{info['electron_count_normalized']:.6f} e-\n"))
        out.append(" --------------------------------------------------\n")
        self._w("".join(out))

    def timings(self, entries) -> None:
        if not entries:
            return
        out = ["\n Setup timings [sec]\n", " --------------------------------------------------\n"]
        out += [f" {label:<35} {t:12.6f}\n" for label, t in entries]
        self._w("".join(out))

    def density_range(self, rho_bohr3_max, rho_bohr3_min) -> None:
        """The line the reference's XC routine logs for every density it is
        given (src/V_xc/exc_nspn.py:43-58)."""
        self._w(f" max and min values of charge density [e/bohr^3]   {rho_bohr3_max:.5e}   {rho_bohr3_min:.5e}\n")

    def initial_density(self, rho_bohr3_max, rho_bohr3_min, hpsum0_ev, exc_ry) -> None:
        self.density_range(rho_bohr3_max, rho_bohr3_min)
        self._w(f" Initial Hartree energy (eV) = {hpsum0_ev:10.5f}  \n"
                f" Initial Exchange-corr. energy (eV) = {exc_ry * RY_EV:10.5f}  \n")

    def start_loop(self) -> None:
        self._w("\n----------------------------------\n\n")

    @staticmethod
    def _energy_lines(e: dict, n_atoms: int, indent: str) -> list[str]:
        rows = [("Sum of eigenvalues     ", e["eigen_sum_ry"]), ("Hartree energy         ", e["hartree_ry"]),
                ("Exchange-corr. energy  ", e["xc_ry"]), ("Ion-ion repulsion      ", e["ion_ry"]),
                ("Total electronic energy", e["total_ry"])]
        out = [f"{indent}{name} = {v * RY_EV:10.5f}  eV   = {v:10.5f}  Ry  \n" for name, v in rows]
        t = e["total_ry"]
        out.append(f"{indent}Electronic energy/atom  = {(t * RY_EV) / n_atoms:10.5f}  eV   = {t / n_atoms:10.5f}  Ry  \n")
        return out

    def iteration(self, it, diag_label, diag_time, lam, occ, hart_tol, hart_label, hart_time, err, energies,
                  n_atoms) -> None:
        """lam in Ha (the operator's eigenvalues), occ in [0, 1]."""
        out = [f"\n\n SCF iter # {it}  ... \n", f"Diagonalization method :\t{diag_label}\n",
               f"Diagonalization time [sec] :\t{diag_time}\n\n",
               "   State  Eigenvalue [Ry]     Eigenvalue [eV]  Occupation\n\n"]
        for i in range(len(occ)):
            ev = float(lam[i]) * 2 * RY_EV
            out.append(f"{i + 1:5d}   {ev / RY_EV:15.10f}   {ev:18.10f}  {2 * float(occ[i]):5.2f}\n")
        out.append(f"Hartree CG tol :\t{hart_tol:.1e} ({hart_label})\n")
        out.append(f"\nHartree potential time [sec]: \t{hart_time}\n\n")
        out.append(f"   ... SCF error = {err:10.2e}\n")
        out.append("   Energy components this iter:\n")
        out += self._energy_lines(energies, n_atoms, "     ")
        self._w("".join(out))

    def summary(self, energies, n_atoms) -> None:
        self._w("".join(["\n\n", " Total Energies \n\n"] + self._energy_lines(energies, n_atoms, " ")))


_ITER = re.compile(r"^\s*SCF iter #\s*(\d+)")
_STATE = re.compile(r"^\s*(\d+)\s+(-?\d+\.\d+)\s+(-?\d+\.\d+)\s+(\d+\.\d+)\s*$")
_ENERGY = re.compile(r"^\s*(Sum of eigenvalues|Hartree energy|Exchange-corr\. energy|Ion-ion repulsion|"
                     r"Total electronic energy|Electronic energy/atom)\s*=\s*(-?\d+\.\d+)\s+eV\s+=\s*(-?\d+\.\d+)\s+Ry")
_ERR = re.compile(r"SCF error =\s*([-+0-9.eE]+)")


def parse_out(path: str) -> dict:
    """Iterations (eigenvalues in Ry, occupations, SCF error, energy components
    in eV and Ry) and the final summary of a log in the reference's format."""
    iters, cur, final, in_final = [], None, {}, False
    with open(path, encoding="utf-8", errors="replace") as f:
        for line in f:
            m = _ITER.match(line)
            if m:
                cur = {"iter": int(m.group(1)), "eig_ry": [], "eig_ev": [], "occ": [], "energy": {}}
                iters.append(cur)
                continue
            if "Total Energies" in line:
                in_final, cur = True, None
                continue
            m = _STATE.match(line)
            if m and cur is not None:
                cur["eig_ry"].append(float(m.group(2)))
                cur["eig_ev"].append(float(m.group(3)))
                cur["occ"].append(float(m.group(4)))
                continue
            m = _ERR.search(line)
            if m and cur is not None:
                cur["err"] = float(m.group(1))
                continue
            m = _ENERGY.match(line)
            if m:
                tgt = final if in_final else (cur["energy"] if cur is not None else None)
                if tgt is not None:
                    tgt[m.group(1)] = (float(m.group(2)), float(m.group(3)))
    return {"iterations": iters, "final": final}


# ---------------------------------------------------------------------------- densities
def _cube_flat(values_domain, grid) -> np.ndarray:
    v = np.asarray(values_domain, dtype=np.float64).reshape(-1)
    if v.size == grid.nx * grid.ny * grid.nz:
        return v
    if v.size != grid.npts:
        raise ValueError(f"density of {v.size} values does not match the domain ({grid.npts} points)")
    return grid.scatter_to_cube(v)


def save_density_variants(rho_cell, grid, base_path: str):
    """rho_cell: electrons per cell on the domain (or the whole cube).  Returns
    the three electron counts (grid, bohr^-3, Angstrom^-3 sums)."""
    flat = _cube_flat(rho_cell, grid)
    h = float(grid.h)
    g3 = flat.reshape((grid.nx, grid.ny, grid.nz), order="F")
    b3 = g3 / h ** 3
    a3 = b3 / A0_ANG ** 3
    d = os.path.dirname(base_path)
    if d:
        os.makedirs(d, exist_ok=True)
    np.save(f"{base_path}_grid.npy", g3)
    np.save(f"{base_path}_bohr3.npy", b3)
    np.save(f"{base_path}.npy", a3)
    h_ang = h * A0_ANG
    return float(g3.sum()), float(b3.sum() * h ** 3), float(a3.sum() * h_ang ** 3)


def load_density_ang3(path: str, grid) -> np.ndarray:
    """An (nx, ny, nz) Angstrom^-3 density file -> electrons per cell on the
    domain points (before clipping and normalisation)."""
    a3 = np.load(path, allow_pickle=False)
    if a3.ndim != 3 or a3.shape != (grid.nx, grid.ny, grid.nz):
        raise ValueError(f"density file shape {a3.shape} does not match the grid {(grid.nx, grid.ny, grid.nz)}")
    flat = np.ravel(np.asarray(a3, dtype=np.float64) * A0_ANG ** 3, order="F") * grid.h ** 3
    return grid.gather_from_cube(flat)


def density_start_from_file(prob, path: str, cg_maxit: int | None = None, cg_tol: float = 1e-5):
    """(rho1, V_H1, info) for run_scf(rho_init=rho1, hpot_init=V_H1) from a
    saved Angstrom^-3 density (the reference's ML-density start)."""
    import torch

    from .hartree import HartreeSolver
    raw = load_density_ang3(path, prob.grid)
    count_raw = float(raw.sum())
    rho1 = np.maximum(raw, 0.0)
    rho1 = prob.nelec * rho1 / float(rho1.sum())
    rho1_t = torch.tensor(rho1, dtype=torch.float64, device="cuda")
    h3 = prob.h ** 3
    rhs = (4.0 * math.pi / h3) * (rho1_t - prob.rho0)
    dv = torch.zeros_like(rhs)
    if cg_maxit is None:
        cg_maxit = int(getattr(prob.cfg, "pcg_maxit", 200))
    hs = HartreeSolver(prob.H.dgrid, prob.h, reference_aliasing=False, prec=getattr(prob.cfg, "pcg_prec", "none"))
    its = hs.solve(rhs, dv, cg_maxit, cg_tol)
    hs.close()
    info = {"electron_count_initial": count_raw, "electron_count_normalized": float(rho1.sum()),
            "electron_target": float(prob.nelec), "negative_points": int(np.sum(raw < 0)), "cg_its": its}
    return rho1_t, prob.hpot0 + dv, info


# ---------------------------------------------------------------------------- wfn.dat
def save_wavefunction(path: str, grid, potential, rho, W, nev: int, atoms) -> None:
    """potential, rho: domain vectors (or whole cube); W: (>= nev, >= npts)
    rows = states (device or host); atoms: [(symbol, (n, 3) bohr)]."""
    pot = _cube_flat(_host(potential), grid)
    rh = _cube_flat(_host(rho), grid)
    Wh = np.asarray(_host(W), dtype=np.float64)
    with open(path, "wb") as f:
        f.write(np.array(26, dtype=np.uint32).tobytes())
        f.write(np.array(grid.rad, dtype=np.float64).tobytes())
        f.write(np.array(grid.h, dtype=np.float64).tobytes())
        f.write(np.array(pot.size, dtype=np.uint32).tobytes())
        f.write(pot.astype(np.float64).tobytes())
        f.write(np.array(rh.size, dtype=np.uint32).tobytes())
        f.write(rh.astype(np.float64).tobytes())
        n = grid.nx * grid.ny * grid.nz
        f.write(np.array(n, dtype=np.uint32).tobytes())
        f.write(np.array(nev, dtype=np.uint32).tobytes())
        for s in range(nev):
            f.write(_cube_flat(Wh[s, :grid.npts], grid).astype(np.float64).tobytes())
        f.write(np.array(len(atoms), dtype=np.uint32).tobytes())
        for _, xyz in atoms:
            xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
            f.write(np.array(xyz.shape[0], dtype=np.uint32).tobytes())
            for r in xyz:
                f.write(r.tobytes())


def read_wavefunction(path: str) -> dict:
    with open(path, "rb") as f:
        def u32():
            return int(np.frombuffer(f.read(4), dtype=np.uint32)[0])

        def f64(k):
            return np.frombuffer(f.read(8 * k), dtype=np.float64).copy()
        tag = u32()
        radius, h = f64(1)[0], f64(1)[0]
        pot = f64(u32())
        rho = f64(u32())
        n, nev = u32(), u32()
        W = np.stack([f64(n) for _ in range(nev)]) if nev else np.zeros((0, n))
        atoms = []
        for _ in range(u32()):
            cnt = u32()
            atoms.append(f64(3 * cnt).reshape(cnt, 3))
    return {"tag": tag, "radius": radius, "h": h, "potential": pot, "rho": rho, "W": W, "atoms": atoms}


def _host(x):
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return x.detach().double().cpu().numpy()
    except ImportError:
        pass
    return np.asarray(x)
