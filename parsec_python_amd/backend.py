"""Drop-in GPU solver backend with the reference's ``SolverBackend`` interface.

``select_solver_backend(settings)`` plays the part of
``src/rsdft_backend.py:80-137``: with ``settings.use_gpu`` true it returns a
``SolverBackend`` bundle (the dataclass of ``src/rsdft_models.py:179-217``,
same slot names) whose stages run on this package's HIP library, so the
reference's SCF driver (``src/rsdft_solver.py:281-395``) can run unchanged:

* ``pseudo_diag(domain, atoms, elem, n_elements, return_info=True)`` ->
  ``(rho0, hpot0, ppot, info)`` as NumPy arrays on the domain's cube grid
  (``src/rsdft_solver.py:87``); it also creates the device context (grid
  map, projector tables) the later stages use.
* ``pseudo_nl(domain, atoms, elem, n_elements)`` -> a ``NonlocalOperator``.
  The driver forms ``B = (0.5*A + vnl) + sp.diags(pot)``
  (``src/rsdft_solver.py:263,285``); with this operator the sums produce a
  ``KohnShamOperator`` that keeps the device stencil and the factored
  projectors and takes ``pot`` from the diagonal matrix -- nothing is
  assembled as CSR.  ``A`` (the SciPy Laplacian the driver built) is checked
  against the device stencil (shape, radius, weights) when it is added.
* ``first_filt(k, B, polm)`` / ``chefsi1(W, ritz, deg, nev, B)`` -> ``(W, ritz)``
  with ``W`` an N x k torch view of the device block (``W[:, :nev] ** 2 @ occ``
  works on it) and ``ritz`` a NumPy vector.
* ``occupations(lam, T, Ne, tol)`` -> ``(E_F, occ)``, ``occ`` a device tensor of
  the W block's dtype.
* ``pcg(A, rhs, x0, maxit, tol, *prec)`` -> ``(x, its)`` on the device; the
  preconditioner arguments are ignored with a one-time warning, as the
  reference's GPU twin does (``src/Eigensolvers/pcg_gpu.py:40-44``).
* ``lanczos(B, nev, v, m, tol)`` and ``chsubsp(deg, nev, B)`` (diagmeth 0/1/2)
  run the device Lanczos of ``Eigensolver.lanczos`` / ``Eigensolver.chsubsp``
  and return ``(W, ritz)`` like ``first_filt``.

``to_numpy_array`` / ``to_host_scalar`` accept torch tensors.  Failures of the
HIP library raise ``ParsecHipError``; there is no CPU fallback.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Any, Callable

import numpy as np
import scipy.sparse as sp
import torch

from . import ionic
from .device import F32, F64, require_gpu
from .eigensolvers import Eigensolver
from .grid import DeviceGrid, fd_weights, make_grid
from .hamiltonian import Hamiltonian
from .hartree import HartreeSolver
from .occupations import occupations as _occupations_host


@dataclass
class SolverBackend:
    """Implementation bundle for the solver stages (``src/rsdft_models.py:179-217``)."""
    label: str
    pseudo_diag: Callable[..., Any]
    pseudo_nl: Callable[..., Any]
    first_filt: Callable[..., Any]
    chefsi1: Callable[..., Any]
    lanczos: Callable[..., Any]
    chsubsp: Callable[..., Any]
    occupations: Callable[..., Any]
    pcg: Callable[..., Any]
    array_module: Any
    cupy_module: Any | None = None

    def to_numpy_array(self, value: Any) -> Any:
        if isinstance(value, torch.Tensor):
            return value.detach().cpu().numpy()
        return value

    def to_host_scalar(self, value: Any) -> Any:
        if isinstance(value, torch.Tensor):
            return value.detach().cpu().reshape(-1)[0].item() if value.numel() == 1 else value.detach().cpu().numpy()
        if isinstance(value, np.ndarray) and value.ndim == 0:
            return value.item()
        if isinstance(value, np.generic):
            return value.item()
        return value


def _radius_of(A) -> int:
    """Stencil radius of a SciPy FD Laplacian from its widest row (6R+1 entries)."""
    A = sp.csr_matrix(A)
    nnz = int(np.max(np.diff(A.indptr))) if A.shape[0] else 0
    if (nnz - 1) % 6:
        raise ValueError("A is not a 3D central-difference Laplacian")
    return (nnz - 1) // 6


class GpuContext:
    """Device state shared by the backend stages of one run (one grid, one atom set)."""

    def __init__(self, domain: dict, atoms, precision: str = "fp64", seed: int = 0, pcg_aliasing: bool = True,
                 pcg_prec: str = "none"):
        require_gpu()
        nx, ny, nz = int(domain["nx"]), int(domain["ny"]), int(domain["nz"])
        if not (nx == ny == nz):
            raise ValueError("the domain must be a cube (nx == ny == nz)")
        self.domain = dict(domain)
        self.grid = make_grid(nx, float(domain["h"]), float(domain["radius"]), "cube")
        self.n = self.grid.npts
        self.h = self.grid.h
        self.atoms = ionic.normalize_atoms(atoms)
        self.precision = precision
        self.seed = seed
        self.pcg_aliasing = pcg_aliasing
        self.pcg_prec = pcg_prec
        self.radius = None
        self.H = None
        self.eig = None
        self.hart = None
        self.proj = None
        self._pcg_grid = {}
        self.w_dtype = F64

    # Hamiltonian pieces, built when the driver first adds 0.5*A to V_nl
    def ensure_operator(self, radius: int):
        if self.H is not None:
            if radius != self.radius:
                raise ValueError("Laplacian stencil radius changed during the run")
            return
        self.radius = radius
        self.proj = ionic.DeviceProjectors(self.grid, self.atoms)
        dtype = F64 if self.precision == "fp64" else F32
        self.H = Hamiltonian(self.grid, radius, self.proj, torch.zeros(self.n, dtype=F64, device="cuda"), dtype)

    def eigensolver(self, k: int) -> Eigensolver:
        if self.eig is None or self.eig.kmax < k:
            if self.eig is not None:
                self.eig.close()
            self.eig = Eigensolver(self.H, k, seed=self.seed)
        return self.eig

    def hartree(self, radius: int) -> HartreeSolver:
        if radius not in self._pcg_grid:
            dg = self.H.dgrid if (self.H is not None and self.radius == radius) else DeviceGrid(self.grid, radius)
            self._pcg_grid[radius] = HartreeSolver(dg, self.h, reference_aliasing=self.pcg_aliasing,
                                                   prec=self.pcg_prec)
        return self._pcg_grid[radius]

    def check_laplacian(self, A, scale: float) -> int:
        """Check that ``A`` (times ``scale``) is the device stencil's -lap/h^2; return its radius."""
        if not sp.issparse(A) or A.shape != (self.n, self.n):
            raise ValueError(f"expected the {self.n} x {self.n} SciPy FD Laplacian of the domain")
        radius = _radius_of(A)
        w = [float(c) for c in fd_weights(radius)]
        diag = A.diagonal() / scale
        want = 3.0 * w[0] / (self.h * self.h)
        if not np.allclose(diag, want, rtol=1e-12, atol=0.0):
            raise ValueError("A's diagonal does not match the central-difference stencil of this order")
        # one interior off-diagonal per axis
        c = self.n // 2
        row = sp.csr_matrix(A).getrow(c).toarray().ravel() / scale
        for d in range(1, radius + 1):
            for step in (d, d * self.grid.nx, d * self.grid.nx * self.grid.ny):
                if c + step < self.n and row[c + step] != 0.0:
                    if not math.isclose(row[c + step], w[d] / (self.h * self.h), rel_tol=1e-12):
                        raise ValueError("A's off-diagonal weights do not match the stencil")
        return radius


class NonlocalOperator:
    """V_nl of the domain (factored KB projectors on the device)."""

    __array_priority__ = 100   # make SciPy defer ``sparse + NonlocalOperator`` to __radd__

    def __init__(self, ctx: GpuContext):
        self.ctx = ctx
        self.shape = (ctx.n, ctx.n)

    def __radd__(self, other):
        return self.__add__(other)

    def __add__(self, other):
        # other = 0.5 * A with A the driver's Laplacian
        radius = self.ctx.check_laplacian(other, 0.5)
        self.ctx.ensure_operator(radius)
        return KohnShamOperator(self.ctx, None)


class KohnShamOperator:
    """0.5*A + V_nl (+ diag(pot) once a diagonal matrix is added)."""

    __array_priority__ = 100

    def __init__(self, ctx: GpuContext, pot):
        self.ctx = ctx
        self.pot = pot
        self.shape = (ctx.n, ctx.n)

    def __add__(self, other):
        if self.pot is not None:
            raise ValueError("the potential has already been added")
        if not sp.issparse(other) or other.shape != self.shape:
            raise ValueError("only sp.diags(pot) can be added to the Kohn-Sham operator")
        dia = sp.dia_matrix(other)
        if np.any(dia.offsets != 0):
            raise ValueError("only a diagonal matrix can be added to the Kohn-Sham operator")
        pot = np.asarray(other.diagonal(), dtype=np.float64)
        return KohnShamOperator(self.ctx, pot)

    __radd__ = __add__

    def bind(self):
        if self.pot is None:
            raise ValueError("the Kohn-Sham operator has no potential (add sp.diags(pot))")
        self.ctx.H.set_potential(torch.from_numpy(self.pot).to("cuda"))
        return self.ctx.H

    def __matmul__(self, X):
        """H @ X for an N x m NumPy array or torch tensor (result on the device, FP64)."""
        H = self.bind()
        Xt = torch.as_tensor(np.asarray(X) if not isinstance(X, torch.Tensor) else X, device="cuda",
                             dtype=F64)
        vec = Xt.dim() == 1
        Xt = Xt.reshape(self.ctx.n, -1)
        m = Xt.shape[1]
        Yb = torch.zeros((m, H.ld), dtype=F64, device="cuda")
        Yb[:, :self.ctx.n] = Xt.T
        Ob = torch.empty_like(Yb)
        dt = H.dtype
        H.dtype = F64
        try:
            H.apply(Yb, Ob, m)
        finally:
            H.dtype = dt
        out = Ob[:, :self.ctx.n].T
        return out.reshape(-1) if vec else out


class GpuBackendImpl:
    """The stage functions, bound to one GpuContext."""

    def __init__(self, precision: str = "fp64", seed: int = 0, pcg_aliasing: bool = True, pcg_prec: str = "none"):
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if pcg_prec not in ("none", "mg"):
            raise ValueError("pcg_prec must be 'none' or 'mg'")
        self.pcg_prec = pcg_prec
        self.precision = precision
        self.seed = seed
        self.pcg_aliasing = pcg_aliasing
        self.ctx: GpuContext | None = None
        self.warned_preconditioner = False

    def _context(self, domain, atoms) -> GpuContext:
        if self.ctx is None:
            self.ctx = GpuContext(domain, atoms, self.precision, self.seed, self.pcg_aliasing, self.pcg_prec)
        return self.ctx

    # -- src/V_ion/pseudoDiag.py:44-421
    def pseudo_diag(self, domain, atoms, elem=None, n_elements=None, return_info=False):
        ctx = self._context(domain, atoms)
        dg = DeviceGrid(ctx.grid, 1)
        try:
            rho0, hpot0, pot, info = ionic.pseudo_diag(ctx.grid, dg, ctx.atoms)
            out = (rho0.cpu().numpy(), hpot0.cpu().numpy(), pot.cpu().numpy())
        finally:
            dg.close()
        return out + (info,) if return_info else out

    # -- src/V_ion/pseudoNL_original.py:9-208
    def pseudo_nl(self, domain, atoms, elem=None, n_elements=None):
        return NonlocalOperator(self._context(domain, atoms))

    def _operator(self, B) -> Hamiltonian:
        if not isinstance(B, KohnShamOperator) or B.ctx is not self.ctx:
            raise ValueError("B must be 0.5*A + pseudo_nl(...) + sp.diags(pot) built from this backend")
        H = B.bind()
        return H

    def _w_view(self, W):
        n = self.ctx.n
        self.ctx.w_dtype = W.dtype
        return W[:, :n].T

    # -- src/Eigensolvers/first_filt.py:6-75
    def first_filt(self, nev, B, polm):
        H = self._operator(B)
        es = self.ctx.eigensolver(nev)
        W, ritz = es.first_filt(nev, polm)
        return self._w_view(W), np.asarray(ritz, dtype=np.float64)

    # -- src/Eigensolvers/chefsi1.py:10-69
    def chefsi1(self, Vin, ritzv, deg, nev, B):
        H = self._operator(B)
        ritzv = np.asarray(self.to_numpy(ritzv), dtype=np.float64).reshape(-1)
        k = len(ritzv)
        es = self.ctx.eigensolver(k)
        block = None
        if isinstance(Vin, torch.Tensor) and Vin.is_cuda:
            for b in es.bufs:
                if Vin.data_ptr() == b.data_ptr() and Vin.dtype == b.dtype:
                    block = b
                    break
        if block is None:   # a foreign N x k array: copy it into a block
            block = es.bufs[0]
            Xt = torch.as_tensor(self.to_numpy(Vin) if not isinstance(Vin, torch.Tensor) else Vin, device="cuda")
            block[:k, :self.ctx.n].copy_(Xt.reshape(self.ctx.n, k).T.to(block.dtype))
        W, ritz = es.chefsi1(block, ritzv, deg, nev)
        return self._w_view(W), np.asarray(ritz, dtype=np.float64)

    # -- src/Eigensolvers/occupations.py:5-56
    def occupations(self, lam, temp, nelec, tol):
        ef, occ = _occupations_host(np.asarray(self.to_numpy(lam), dtype=np.float64), temp, nelec, tol)
        dtype = self.ctx.w_dtype if self.ctx is not None else F64
        return ef, torch.as_tensor(np.asarray(occ, dtype=np.float64), device="cuda").to(dtype)

    # -- src/Eigensolvers/pcg.py:7-85 / pcg_gpu.py:22-65
    def pcg(self, A, rhs, x0, m, tol, *prec):
        if len(prec) >= 2 and not self.warned_preconditioner:
            print("Warning: GPU PCG ignores the CPU-side preconditioner arguments.")
            self.warned_preconditioner = True
        if self.ctx is None:
            raise ValueError("pcg: call pseudo_diag first (it creates the device context)")
        radius = self.ctx.check_laplacian(A, 1.0)
        hs = self.ctx.hartree(radius)
        b = torch.as_tensor(np.asarray(self.to_numpy(rhs), dtype=np.float64).reshape(-1), device="cuda")
        x = torch.as_tensor(np.asarray(self.to_numpy(x0), dtype=np.float64).reshape(-1), device="cuda").clone()
        if b.numel() != self.ctx.n or x.numel() != self.ctx.n:
            raise ValueError("rhs and x0 must have one entry per grid point")
        its = hs.solve(b, x, int(m), float(tol))
        return x, its

    @staticmethod
    def to_numpy(v):
        if isinstance(v, torch.Tensor):
            return v.detach().cpu().numpy()
        return v

    # -- src/Eigensolvers/lanczos.py:9-223 (diagmeth 0 at the first iteration, 1 at every one)
    def lanczos(self, B, nev, v, m, tol, *progress):
        """nev Ritz pairs from m Lanczos steps; ``v`` (the reference's random
        start vector) is not used: the start vector comes from the backend's
        seeded counter-based generator."""
        H = self._operator(B)
        es = self.ctx.eigensolver(nev)
        W, ritz = es.lanczos(int(nev), int(m), float(tol))
        return self._w_view(W), np.asarray(ritz, dtype=np.float64)

    # -- src/Eigensolvers/chsubsp.py:10-91 (diagmeth 2, first iteration)
    def chsubsp(self, deg, nev, B):
        H = self._operator(B)
        es = self.ctx.eigensolver(nev)
        W, ritz = es.chsubsp(int(nev), int(deg))
        return self._w_view(W), np.asarray(ritz, dtype=np.float64)


def make_backend(precision: str = "fp64", seed: int = 0, pcg_aliasing: bool = True,
                 pcg_prec: str = "none") -> SolverBackend:
    impl = GpuBackendImpl(precision, seed, pcg_aliasing, pcg_prec)
    return SolverBackend(label="gpu", pseudo_diag=impl.pseudo_diag, pseudo_nl=impl.pseudo_nl,
                         first_filt=impl.first_filt, chefsi1=impl.chefsi1, lanczos=impl.lanczos,
                         chsubsp=impl.chsubsp, occupations=impl.occupations, pcg=impl.pcg, array_module=torch,
                         cupy_module=None)


def select_solver_backend(settings) -> SolverBackend:
    """``use_gpu`` true -> the HIP backend; the CPU bundle stays the reference's own.

    Optional settings attributes: ``gpu_precision`` ("fp64" | "mixed"),
    ``gpu_seed`` (start-vector seed), ``gpu_pcg_aliasing`` (reproduce the CPU
    CG's first-iteration aliasing; default True), ``gpu_pcg_prec`` ("none",
    the reference's default unpreconditioned CG, or "mg": multigrid-
    preconditioned CG with the same stop rule).
    """
    if not getattr(settings, "use_gpu", 0):
        raise ValueError("this backend is the GPU path (use_gpu=1); the CPU path is the reference's own bundle")
    try:
        from . import _lib
        _lib.lib()
    except Exception as e:  # the library must be present: no silent fallback
        raise SystemExit(f"GPU backend requested but the HIP library could not be loaded: {e}")
    require_gpu()
    return make_backend(getattr(settings, "gpu_precision", "fp64"), getattr(settings, "gpu_seed", 0),
                        getattr(settings, "gpu_pcg_aliasing", True), getattr(settings, "gpu_pcg_prec", "none"))
