"""Run a PARSEC.py input file on the GPU, with the reference's outputs.

``run_input_file(path)`` is the counterpart of ``src/main_new.py:46-111`` plus
``run_rsdft_calculation`` (``src/rsdft_solver.py:156-462``) for an input file:

* the file is read by ``inputs.read_input`` (``.in``/``.inp``, ``.json``,
  ``.txt``, ``.dat``, ``.mat``), lengths converted to bohr and the solver
  settings (tol, maxits, fd_order, Fermi_temp, poldeg, diagmeth, CG_prec,
  nev, grid spacing, radius, recenter_atoms, save_wfn, mixer) applied;
* the atoms are recentred for a SAD start (``src/rsdft_setup.py:288-295``),
  the grid sized as the reference sizes it, the SCF run on the device with
  the chosen eigensolver schedule (diagmeth) and mixer;
* an ML-density run (``density_method = ml``) starts from the Angstrom^-3
  density file named by ``ml_file_path``; its array must have the grid's shape
  (the reference sizes such a grid from a POSCAR instead, which is not read
  here), and without a diagmeth in the file diagmeth 2 is used, as in the
  reference (``src/main_new.py:79-81``);
* the outputs go next to the input file, named after it
  (``src/rsdft_setup.py:215-249``): ``<base>.out`` (the reference's log
  layout), ``<base>_init_rho{,_bohr3,_grid}.npy``,
  ``<base>_conv_rho{,_bohr3,_grid}.npy`` and, with ``save_wfn = 1``,
  ``<base>_wfn.dat``.

The reference's domain is the full cube; ``domain="sphere"`` selects the
sphere-masked grid of the same size instead.
"""
from __future__ import annotations

import os
import time
from dataclasses import dataclass

from . import molecules as M
from .inputs import read_input, solver_settings, to_bohr
from .outputs import OutLog, RY_EV, density_start_from_file, save_density_variants, save_wavefunction


@dataclass
class RunOutputs:
    output_file: str
    init_density_base: str
    conv_density_base: str
    wfn_file: str | None


def output_paths(input_path: str, out_dir: str | None = None) -> RunOutputs:
    base = os.path.splitext(os.path.basename(input_path))[0]
    d = out_dir or os.path.dirname(os.path.abspath(input_path))
    stem = os.path.join(d, base)
    return RunOutputs(f"{stem}.out", f"{stem}_init_rho", f"{stem}_conv_rho", f"{stem}_wfn.dat")


def _xc_energy(prob, rho_bohr3) -> float:
    """Exchange-correlation energy (Ry) of a density in e/bohr^3, on the device."""
    import ctypes as C

    import torch

    from . import _lib
    from .device import ptr, stream
    n = prob.n
    nb = C.c_int()
    _lib.call("pc_num_blocks", n, C.byref(nb))
    dev = dict(dtype=torch.float64, device="cuda")
    part = torch.empty(5 * nb.value, **dev)
    sums = torch.empty(8, **dev)
    vxc = torch.empty(n, **dev)
    _lib.call("pc_xc", n, ptr(rho_bohr3), ptr(vxc), ptr(part), ptr(sums), stream())
    return float(sums[0]) * prob.h ** 3


def run_input_file(path: str, out_dir: str | None = None, precision: str = "fp64", domain: str = "cube",
                   seed: int = 0, pcg_prec: str | None = None, write: bool = True, **cfg_overrides):
    """Run the input file; returns (SCFResult, SCFProblem, RunOutputs).  The
    problem stays open (``prob.close()`` frees its device memory)."""
    import torch

    from .scf import SCFConfig, SCFProblem, run_scf
    inp = read_input(path)
    atoms, over = to_bohr(inp)
    st = solver_settings(over)
    has_diag = any(str(k).lower() == "diagmeth" for k in inp.settings)
    if inp.density_method in ("ml", "sad_ml_grid") and not has_diag:
        st["diagmeth"] = 2
    if inp.density_method == "sad" and st["recenter_atoms"]:
        atoms = M.recenter(atoms)
    dom, nev, zelec = M.size_domain(atoms, st["h"], st["radius"], st["nev"])
    cfg = SCFConfig(fd_order=st["fd_order"], maxits=st["maxits"], tol=st["tol"], fermi_temp=st["fermi_temp"],
                    poldeg=st["poldeg"], precision=precision, domain=domain, seed=seed, diagmeth=st["diagmeth"],
                    mixer=st["mixer"], z_charge=inp.z_charge,
                    pcg_prec=pcg_prec if pcg_prec is not None else ("mg" if st["cg_prec"] else "none"))
    for k, v in cfg_overrides.items():
        setattr(cfg, k, v)
    paths = output_paths(path, out_dir)
    t_setup = time.perf_counter()
    prob = SCFProblem(atoms, dom, nev, cfg)
    t_setup = time.perf_counter() - t_setup
    log = OutLog(paths.output_file, "gpu", inp.density_method) if write else None
    if log:
        log.parameters(nev, prob.atoms, dom, cfg.poldeg, cfg.fd_order)
    rho_init = hpot_init = None
    info = dict(prob.diag_info)
    if inp.density_method in ("ml", "sad_ml_grid"):
        if not inp.ml_file_path:
            raise SystemExit("Density method requires ML .npy file but no ml_file_path was provided.")
        rho_init, hpot_init, info = density_start_from_file(prob, inp.ml_file_path)
    rho_start = prob.rho0 if rho_init is None else rho_init
    hpot_start = prob.hpot0 if hpot_init is None else hpot_init
    h = prob.h
    if log:
        log.electron_counts(info)
        save_density_variants(rho_start.cpu().numpy(), prob.grid, paths.init_density_base)
        rho_b3 = rho_start / h ** 3
        exc_ry = _xc_energy(prob, rho_b3)
        log.initial_density(float(rho_b3.max()), float(rho_b3.min()),
                            float(torch.dot(rho_start, hpot_start)) * RY_EV, exc_ry)
        log.timings([("Setup (grid, ionic potentials, projectors)", t_setup)])
    hart_label = "multigrid prec" if cfg.pcg_prec == "mg" else "no prec"

    if log:
        log.start_loop()

    def on_iter(r):
        if log:
            log.density_range(float(r["rho"].max()), float(r["rho"].min()))
            log.iteration(r["iter"], r["diag_label"], r["diag_time"], r["lam"], r["occ"], cfg.pcg_tol, hart_label,
                          r["hart_time"], r["err"], r["energies"], prob.n_atoms)

    res = run_scf(prob, rho_init=rho_init, hpot_init=hpot_init, callback=on_iter)
    if log:
        log.summary(res.energies, prob.n_atoms)
        save_density_variants((res.rho * h ** 3).cpu().numpy(), prob.grid, paths.conv_density_base)
        if st["save_wfn"]:
            save_wavefunction(paths.wfn_file, prob.grid, res.pot, res.rho, res.W, nev, prob.atoms)
    if not st["save_wfn"]:
        paths.wfn_file = None
    torch.cuda.synchronize()
    return res, prob, paths
