"""Hartree potential: CG solve of (-lap/h^2) V_H = 4 pi (rho - rho0)/h^3 on the device.

Mirrors ``pcg(A, rhs, x0, m, tol)`` (``src/Eigensolvers/pcg.py:7-85``; called at
``src/rsdft_solver.py:345`` with m=200, tol=1e-5, warm start from the previous
V_H).  FP64 throughout.  ``reference_aliasing=True`` reproduces the CPU
reference's first-iteration p/r aliasing (its second search direction is
(1+beta) r_1); False gives textbook CG (the reference's CuPy twin).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .device import Scratch, ptr, stream
from .grid import DeviceGrid


class HartreeSolver:
    """``prec="mg"`` selects the multigrid-preconditioned CG (pc_pcg_solve_mg):
    same system and stop rule, fewer iterations; the reference's optional
    preconditioned CG (cg_prec) is the counterpart."""

    def __init__(self, dgrid: DeviceGrid, h: float, reference_aliasing: bool = True, check_every: int = 16,
                 prec: str | None = None):
        self.dgrid = dgrid
        self.inv_h2 = 1.0 / (h * h)
        self.alias = int(bool(reference_aliasing))
        self.check_every = check_every
        self.mg = None
        nb = C.c_size_t()
        if prec == "mg":
            handle = C.c_void_p()
            _lib.call("pc_mg_create", dgrid.handle, h, C.byref(handle))
            self.mg = handle
            _lib.call("pc_mg_workspace", self.mg, C.byref(nb))
        elif prec in (None, "none"):
            _lib.call("pc_pcg_workspace", dgrid.handle, C.byref(nb))
        else:
            raise ValueError("prec must be None, 'none' or 'mg'")
        self.work_bytes = nb.value
        self.scratch = Scratch()
        self.last_its = 0
        self.last_res = 0.0

    def solve(self, rhs: torch.Tensor, x: torch.Tensor, maxit: int = 200, tol: float = 1e-5) -> int:
        """x (warm start, overwritten with the solution); returns iterations."""
        work = self.scratch.bytes("pcg", self.work_bytes)
        its = C.c_int()
        res = (C.c_double * 2)()
        if self.mg is not None:
            # after a first solve, queue about as many iterations as it took
            # before the first readback, then read back every 2: fewer
            # iterations queued past convergence (the results are the same)
            first = max(1, self.last_its - 1) if self.last_its > 0 else 0
            _lib.call("pc_mg_set_first_check", self.mg, first)
            _lib.call("pc_pcg_solve_mg", self.mg, ptr(rhs), ptr(x), maxit, tol, 2 if first else self.check_every,
                      ptr(work), self.work_bytes, C.byref(its), res, stream())
        else:
            _lib.call("pc_pcg_solve", self.dgrid.handle, self.inv_h2, ptr(rhs), ptr(x), maxit, tol, self.alias,
                      self.check_every, ptr(work), self.work_bytes, C.byref(its), res, stream())
        self.last_its = its.value
        self.last_res = res[0]
        # relative residual |r|/|r0| reached (tol^2 r0.r0 was stored)
        self.last_rel = (res[0] * tol * tol / res[1]) ** 0.5 if res[1] > 0 else 0.0
        # initial residual norm (tol^2 * r0.r0 was stored)
        self.last_r0 = (res[1] ** 0.5) / tol if tol > 0 else float("nan")
        return its.value

    def close(self):
        if self.mg is not None:
            _lib.call("pc_mg_destroy", self.mg)
            self.mg = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
