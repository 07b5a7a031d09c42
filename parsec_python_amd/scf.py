"""The SCF driver on one GPU (mirrors ``run_rsdft_calculation``,
``src/rsdft_solver.py:156-462``, diagmeth 3).

Per iteration (``src/rsdft_solver.py:281-395``):
  H <- 0.5 A + V_nl + diag(pot)                      (resident operator, new pot)
  iteration 1: first_filt(nev+15, H, poldeg); later: chefsi1(W, lam, poldeg, nev, H)
  occupations(lam[:nev], T, Nelec, 1e-6)            (host FP64)
  rho = sum_i 2 f_i W_i^2                            (device)
  V_H <- CG(A, 4 pi/h^3 (rho - rho0), warm start)    (device FP64)
  v_xc, pot_new, |pot_new - pot|/|pot_new|, energy sums (one device pass)
  pot <- mixer(pot, pot_new - pot)                   (device)
Energies in Ry as ``compute_energy_components`` (:128-153).  The run stops on
the reference's rule err <= tol, and optionally on |E_i - E_{i-1}| < etol
Ha/atom.

Precision: ``fp64`` runs everything in FP64.  ``mixed`` keeps the wave
functions, the filter, H*X and the subspace GEMMs in FP32 (MFMA, FP32
accumulation within a chunk, FP64 across chunks) and keeps the Cholesky /
eigensolver, density, Hartree CG, mixer and energy sums in FP64.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, ionic
from .device import F32, F64, ptr, require_gpu, stream
from .eigensolvers import Eigensolver
from .grid import Grid, make_grid, reference_fd_radius
from .hamiltonian import Hamiltonian, leading_dim
from .hartree import HartreeSolver
from .mixer import make_mixer
from .occupations import occupations

RY_EV = 13.605698066


@dataclass
class SCFConfig:
    fd_order: int = 8
    fd_radius: int | None = None       # explicit stencil radius (e.g. 6 for order 12)
    maxits: int = 40
    tol: float = 2e-4
    fermi_temp: float = 500.0
    poldeg: int = 10
    precision: str = "fp64"            # "fp64" | "mixed"
    domain: str = "cube"               # "cube" | "sphere"
    seed: int = 0
    etol_ha_per_atom: float | None = None
    pcg_aliasing: bool = True
    pcg_maxit: int = 200
    pcg_tol: float = 1e-5
    # Hartree CG preconditioner: "none" (the reference default, cg_prec = 0) or
    # "mg" (multigrid V-cycle; same stop rule, fewer iterations)
    pcg_prec: str = "none"
    # mixed precision: switch the wave functions to FP64 once |dE| falls
    # below polish_de Ha/atom (or err below polish_err) and finish in FP64.
    # The FP32 iteration's energy noise is ~1e-7 of the total energy (below
    # 1e-7 Ha/atom for Si87H76, ~2e-6 Ha/atom for H2O), so the last
    # iterations run in FP64; the convergence test carries across the switch.
    polish: bool = True
    polish_de: float = 5e-6
    polish_err: float | None = None
    # Chebyshev degree of the FP64 iterations after the switch (None: poldeg).
    # By then the FP32 subspace is converged to FP32 accuracy and the FP64
    # Rayleigh-Ritz is what fixes the energy; on Si87H76 (learned start) degree
    # 2, 4, 6 and 10 all finish in 15 iterations, 3.4e-7, 1.8e-7, 4.3e-8 and
    # 5.2e-7 Ha/atom from the converged FP64 energy, 0.430 / 0.434 / 0.438 /
    # 0.446 s.
    polish_deg: int | None = 4
    # eigensolver of the k x k Rayleigh-Ritz problem: "tridiag" (one-launch
    # Householder tridiagonalization + this library's tridiagonal eigensolver,
    # k <= 512; larger k falls back to "syevd"), "syevd" (rocSOLVER divide and
    # conquer) or "syevj" (Jacobi).  All three return the same eigenpairs to
    # rounding.
    dense_method: str = "tridiag"
    # Cholesky factor of the overlap matrix and its inverse: "blocked" (this
    # library's two-launch blocked factor and inverse, then GEMMs; k <= 320,
    # larger k uses "rocsolver") or "rocsolver" (potrf + trtri + trmm)
    dense_cholesky: str = "blocked"
    # Rayleigh-Ritz: S = Y^T Y and its Cholesky factor on a second stream
    # while H Y and G are formed (the same bits either way)
    rr_overlap: bool = True
    # extra fixed-potential filter passes after first_filt when the SCF starts
    # from a given (learned) density, until the mean occupied Ritz value moves
    # by less than start_warmup_tol Ha (see run_scf).  Off by default: on
    # Si87H76 3 passes save 2 SCF iterations and cost about as much.
    start_warmup: int = 0
    start_warmup_tol: float = 1e-3
    # with a given start density (learned start): first_filt's random start
    # vectors are weighted point-wise by sqrt(rho_init / max rho_init), so the
    # start block already leans to where the electrons are.  On Si87H76 (seeds
    # 1-4) this takes the mixed solve from 17 to 15 SCF iterations; the energy
    # at the stop stays within the spread the unweighted start shows.
    weighted_start: bool = True
    # eigensolver schedule (src/rsdft_solver.py:289-316): 3 = first_filt then
    # chefsi1 (default); 2 = chsubsp then chefsi1; 1 = Lanczos every
    # iteration; 0 = Lanczos then chefsi1
    diagmeth: int = 3
    # potential mixer (src/Mixer/mixer.py): "msecant1" (default), "msecant2", "msecant3"
    mixer: str = "msecant1"
    # system charge: electrons removed (src/rsdft_solver.py:224-226,253-255 scale
    # rho0 and V_H0 by (Z - charge)/Z and take charge off the electron count)
    z_charge: float = 0.0
    verbose: bool = False


@dataclass
class SCFResult:
    energies: dict
    eigenvalues: np.ndarray
    occupations: np.ndarray
    iterations: int
    error: float
    converged: bool
    history: list
    timings: dict
    rho: torch.Tensor            # e/bohr^3 on the domain
    hpot: torch.Tensor
    pot: torch.Tensor
    W: torch.Tensor              # (k, ld) block, rows = states
    grid: Grid
    setup_info: dict = field(default_factory=dict)


class SCFProblem:
    """Everything the SCF loop needs, built once (setup)."""

    def __init__(self, atoms, domain: dict, nev: int, cfg: SCFConfig):
        require_gpu()
        t0 = time.perf_counter()
        self.cfg = cfg
        self.atoms = ionic.normalize_atoms(atoms)
        self.nev = int(nev)
        self.k = self.nev + 15
        self.grid = make_grid(int(domain["nx"]), float(domain["h"]), float(domain["radius"]), cfg.domain)
        self.h = self.grid.h
        self.n = self.grid.npts
        self.radius = cfg.fd_radius if cfg.fd_radius is not None else reference_fd_radius(cfg.fd_order)
        self.dtype = F64 if cfg.precision == "fp64" else F32
        self.proj = ionic.DeviceProjectors(self.grid, self.atoms)
        pot0 = torch.zeros(self.n, dtype=F64, device="cuda")
        self.H = Hamiltonian(self.grid, self.radius, self.proj, pot0, self.dtype)
        self.rho0, self.hpot0, self.ppot, self.diag_info = ionic.pseudo_diag(self.grid, self.H.dgrid, self.atoms)
        self.e_nuc = ionic.nuclear(self.atoms)
        self.nelec = ionic.nelectrons(self.atoms)
        if cfg.z_charge != 0:
            zelec = self.nelec
            if zelec - cfg.z_charge < 0:
                raise ValueError("Problem with charge state. Negative number of electrons.")
            scale = (zelec - cfg.z_charge) / zelec
            self.rho0 = self.rho0 * scale
            self.hpot0 = self.hpot0 * scale
            self.nelec = zelec - cfg.z_charge
        self.n_atoms = sum(x.shape[0] for _, x in self.atoms)
        torch.cuda.synchronize()
        self.setup_time = time.perf_counter() - t0

    def close(self):
        self.H.close()
        self.proj.close()


def _xc(n, rho, vxc, part, out1):
    _lib.call("pc_xc", n, ptr(rho), ptr(vxc), ptr(part), ptr(out1), stream())


def run_scf(prob: SCFProblem, rho_init: torch.Tensor | None = None, hpot_init: torch.Tensor | None = None,
            max_wall: float | None = None, warmup: int | None = None,
            start_weight_rho: torch.Tensor | None = None, callback=None) -> SCFResult:
    """Run the SCF loop.  ``rho_init``/``hpot_init`` replace (rho0, V_H0) for a
    learned start, as the reference's ML density path does
    (``src/V_ion/pseudoDiag_ML4Den_poisson.py:392-540``).

    ``callback(record)``, if given, is called after every iteration with the
    iteration number, diagonalisation label and time, eigenvalues and
    occupations, Hartree CG iterations and time, SCF error and energies (what
    the reference writes to its log per iteration) and the new density
    (``rho``, e/bohr^3, the loop's own buffer: read it inside the callback).

    ``start_weight_rho``: the density whose square root weights first_filt's
    random start vectors (``cfg.weighted_start``); default ``rho_init`` when
    it is given, else no weighting.  Passing ``prob.rho0`` gives the SAD start
    with SAD-weighted start vectors, the control that separates the learned
    density's share of the gain from the weighting's.

    ``warmup`` (default: ``cfg.start_warmup`` when ``rho_init`` is given, else
    0): up to that many extra filter + Rayleigh-Ritz passes at the starting
    potential before the first density update, stopped when the mean occupied
    Ritz value moves by less than ``cfg.start_warmup_tol`` Ha.  A start density
    close to the converged one only shortens the SCF once the subspace has
    converged for its potential: otherwise the first output potentials come
    from unconverged (random-start) vectors and the mixer moves away from the
    good density."""
    cfg = prob.cfg
    n, h, nev, k = prob.n, prob.h, prob.nev, prob.k
    h3 = h ** 3
    H = prob.H
    rho0 = prob.rho0 if rho_init is None else rho_init
    hpot0 = prob.hpot0 if hpot_init is None else hpot_init
    ppot = prob.ppot
    dev = dict(dtype=F64, device="cuda")
    torch.cuda.synchronize()
    t_start = time.perf_counter()

    import ctypes as C
    nbc = C.c_int()
    _lib.call("pc_num_blocks", n, C.byref(nbc))
    nb = nbc.value
    part = torch.empty(5 * nb, **dev)
    sums = torch.empty(8, **dev)
    vxc = torch.empty(n, **dev)
    rhoxc = rho0 / h3
    _xc(n, rhoxc, vxc, part, sums)
    pot = ppot + hpot0 + 0.5 * vxc
    hpot = torch.zeros(n, **dev)
    pot_new = torch.empty(n, **dev)
    rho_cell = torch.empty(n, **dev)
    rho = torch.empty(n, **dev)
    rhs = torch.empty(n, **dev)
    f1 = torch.empty(n, **dev)
    pot_next = torch.empty(n, **dev)
    occ2_dev = torch.empty(k, **dev)
    occ2_pin = torch.empty(k, dtype=torch.float64, pin_memory=True)

    H.dtype = prob.dtype
    eig = Eigensolver(H, k, seed=cfg.seed, overlap=cfg.rr_overlap)
    dense_method = {"syevd": 0, "syevj": 1, "tridiag": 2}[cfg.dense_method]
    tri_mode = {"rocsolver": 0, "blocked": 2}[cfg.dense_cholesky]
    _lib.call("pc_dense_set_method", eig.la.dense, dense_method)
    _lib.call("pc_dense_set_triangular", eig.la.dense, tri_mode)
    hart = HartreeSolver(H.dgrid, h, reference_aliasing=cfg.pcg_aliasing, prec=cfg.pcg_prec)
    mixer = make_mixer(cfg.mixer, n, cfg.maxits + 2)
    if cfg.diagmeth not in (0, 1, 2, 3):
        raise ValueError("diagmeth must be 0, 1, 2 or 3")
    lanczos_steps = nev + 500

    err = 10.0 + cfg.tol
    its = 0
    W = None
    lam = None
    occ = None
    hist = []
    first_filt_info = {}
    warmups = 0
    e_prev = None
    energies = None
    converged = False
    tim = {"diag": 0.0, "hartree": 0.0, "post": 0.0}
    while err > cfg.tol and its < cfg.maxits:
        its += 1
        t0 = time.perf_counter()
        H.set_potential(pot)
        if cfg.diagmeth == 1 or (its == 1 and cfg.diagmeth == 0):
            W, lam = eig.lanczos(k, lanczos_steps, 1e-5)
            diag_label = f"lanczos_gpu (diagmeth={cfg.diagmeth})"
        elif its == 1 and cfg.diagmeth == 2:
            W, lam = eig.chsubsp(k, cfg.poldeg)
            diag_label = "chsubsp_gpu (diagmeth=2)"
        elif its == 1:
            diag_label = "first_filt_gpu (diagmeth=3)"
            sw = None
            wrho = start_weight_rho if start_weight_rho is not None else rho_init
            if wrho is not None and cfg.weighted_start:
                r = torch.clamp(wrho, min=0.0)
                sw = torch.sqrt(r / r.max())
            W, lam = eig.first_filt(k, cfg.poldeg, start_weight=sw)
            first_filt_info = dict(eig.last_first_filt)
            nwarm = (cfg.start_warmup if rho_init is not None else 0) if warmup is None else warmup
            mean_prev = float(np.mean(lam[:nev]))
            for _ in range(nwarm):
                W, lam = eig.chefsi1(W, lam, cfg.poldeg, nev)
                warmups += 1
                mean_now = float(np.mean(lam[:nev]))
                if abs(mean_now - mean_prev) < cfg.start_warmup_tol:
                    break
                mean_prev = mean_now
        else:
            deg = cfg.poldeg
            if cfg.polish_deg is not None and H.dtype == F64 and prob.dtype == F32:
                deg = cfg.polish_deg
            W, lam = eig.chefsi1(W, lam, deg, nev)
            diag_label = f"chebsf_gpu (diagmeth={cfg.diagmeth})"
        t1 = time.perf_counter()
        _, occ = occupations(lam[:nev], cfg.fermi_temp, prob.nelec, 1e-6)
        occ2_pin[:nev] = torch.from_numpy(2.0 * np.asarray(occ, dtype=np.float64))
        occ2_dev[:nev].copy_(occ2_pin[:nev], non_blocking=True)   # (the buffer is reused after a readback)
        sfx = "f64" if W.dtype == F64 else "f32"
        _lib.call(f"pc_density_{sfx}", ptr(W), W.shape[1], n, nev, ptr(occ2_dev), ptr(rho_cell), stream())
        _lib.call("pc_hartree_rhs", n, ptr(rho_cell), ptr(rho0), 4 * math.pi / h3, ptr(rhs), stream())
        cg_its = hart.solve(rhs, hpot, cfg.pcg_maxit, cfg.pcg_tol)
        t2 = time.perf_counter()
        _lib.call("pc_scf_post", n, ptr(rho_cell), h3, ptr(ppot), ptr(hpot), ptr(hpot0), ptr(pot), ptr(rho),
                  ptr(vxc), ptr(pot_new), ptr(part), ptr(sums), stream())
        s = sums[:5].cpu().numpy()
        err = float(math.sqrt(s[1]) / math.sqrt(s[2]))
        esum = float(np.sum(lam[:nev] * occ[:nev]))
        e = {"eigen_sum_ry": 4.0 * esum, "hartree_ry": float(s[3] * h3), "xc_ry": float(s[0] * h3),
             "ion_ry": float(prob.e_nuc)}
        vxc_sum = float(s[4] * h3)
        e["total_ry"] = e["eigen_sum_ry"] - e["hartree_ry"] + e["xc_ry"] - vxc_sum + e["ion_ry"]
        energies = e
        e_ha_atom = 0.5 * e["total_ry"] / prob.n_atoms
        de = None if e_prev is None else abs(e_ha_atom - e_prev)
        e_prev = e_ha_atom
        hist.append({"iter": its, "err": err, "total_ry": e["total_ry"], "dE_ha_atom": de, "cg_its": cg_its,
                     "cg_rel": hart.last_rel, "bounds": getattr(eig, "last_bounds", None),
                     "lam": np.array(lam[:nev])})
        if callback is not None:
            callback({"iter": its, "diag_label": diag_label, "diag_time": t1 - t0, "lam": np.array(lam[:nev]),
                      "occ": np.array(occ[:nev]), "cg_its": cg_its, "hart_time": t2 - t1, "err": err,
                      "energies": dict(e), "precision": "fp64" if H.dtype == F64 else "fp32", "rho": rho})
        # pot <- mixer(pot, pot_new - pot)
        f1.copy_(pot_new)
        _lib.call("pc_lincomb", n, None, n, 0, None, 1.0, ptr(pot), -1.0, ptr(f1), stream())
        mixer(pot, f1, pot_next)
        pot, pot_next = pot_next, pot
        t3 = time.perf_counter()
        tim["diag"] += t1 - t0
        tim["hartree"] += t2 - t1
        tim["post"] += t3 - t2
        if cfg.verbose:
            b = getattr(eig, "last_bounds", None)
            bnorm = float(torch.linalg.norm(rhs).item())
            print(f"  SCF {its:3d}  err {err:9.2e}  E {e['total_ry']:.8f} Ry  dE {de if de is not None else 0:.2e}"
                  f" Ha/atom  CG {cg_its} rel {hart.last_rel:.1e} r0/|b| {hart.last_r0 / bnorm:.1e}"
                  f" r/|b| {hart.last_res ** 0.5 / bnorm:.1e}"
                  + (f"  bounds {b[0]:.4f} {b[1]:.4f} {b[2]:.3f}" if b else ""))
        if H.dtype == F32 and cfg.polish and ((de is not None and de < cfg.polish_de) or
                                              (cfg.polish_err is not None and err < cfg.polish_err)):
            # finish in FP64: convert the current Ritz block and keep iterating
            H.dtype = F64
            eig64 = Eigensolver(H, k, seed=cfg.seed, overlap=cfg.rr_overlap)
            _lib.call("pc_dense_set_method", eig64.la.dense, dense_method)
            _lib.call("pc_dense_set_triangular", eig64.la.dense, tri_mode)
            eig64.seeds = eig.seeds
            _lib.call("pc_convert_f32_to_f64", n, k, ptr(W), W.shape[1], ptr(eig64.bufs[0]), eig64.bufs[0].shape[1],
                      stream())
            W = eig64.bufs[0]
            eig.close()
            eig64.stats["filter_steps"] += eig.stats["filter_steps"]
            eig64.stats["rr"] += eig.stats["rr"]
            eig = eig64
            hist[-1]["switch_to_fp64"] = True
            continue       # the next (FP64) energy is tested against this one
        if cfg.etol_ha_per_atom is not None and de is not None and de < cfg.etol_ha_per_atom:
            converged = True
            break
        if max_wall is not None and time.perf_counter() - t_start > max_wall:
            break
    if not converged:
        converged = err <= cfg.tol
    torch.cuda.synchronize()
    tim["total"] = time.perf_counter() - t_start
    tim["per_iteration"] = tim["total"] / max(its, 1)
    res = SCFResult(energies=energies, eigenvalues=np.asarray(lam), occupations=np.asarray(occ), iterations=its,
                    error=err, converged=converged, history=hist, timings=tim, rho=rho, hpot=hpot, pot=pot, W=W,
                    grid=prob.grid, setup_info={"setup_time": prob.setup_time, **prob.diag_info,
                                                "filter_steps": eig.stats["filter_steps"], "rr": eig.stats["rr"],
                                                "first_filt": first_filt_info, "start_warmup": warmups})
    eig.close()
    return res
