"""The Kohn-Sham Hamiltonian as a structured device operator.

The reference assembles ``B = 0.5*A + V_nl + diag(pot)`` as a SciPy CSR matrix
every SCF iteration (``src/rsdft_solver.py:263,285``) and, on its GPU path,
uploads it as FP32 CSR on every call (``src/Eigensolvers/chefsi1_gpu.py:20-41``).
Here the operator stays resident as (stencil weights + row map, potential
vector, factored projectors) and every application is one fused stencil
launch plus a projector gather/scatter.

Blocks of vectors are torch tensors of shape (k, ld) -- vector v is row v,
point p of vector v at ``v*ld + p`` -- so ``block[:, :n].T`` is the reference's
N x k array ``W``.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .device import F32, F64, Scratch, ptr, stream
from .grid import DeviceGrid, Grid
from .ionic import DeviceProjectors

ZMARCH_MIN_VECTORS = 8


def leading_dim(n: int) -> int:
    return (n + 15) // 16 * 16


class Hamiltonian:
    """H = 0.5 * (-lap) + V_nl + diag(pot) on a domain (Hartree units)."""

    def __init__(self, grid: Grid, radius: int, projectors: DeviceProjectors | None, pot: torch.Tensor,
                 dtype: torch.dtype = F64):
        self.grid = grid
        self.dgrid = DeviceGrid(grid, radius)
        self.proj = projectors
        self.n = grid.npts
        self.ld = leading_dim(self.n)
        self.h = grid.h
        self.dtype = dtype
        self.kin_scale = 0.5 / (grid.h * grid.h)
        self.scratch = Scratch()
        self.pot = None
        self.pot32 = None
        self.set_potential(pot)
        # optional per-launch timing of the multi-vector stencil: list of
        # (start event, end event, algorithmic bytes, element size)
        self.stencil_timer: list | None = None

    # ------------------------------------------------------------------ potential
    def set_potential(self, pot: torch.Tensor) -> None:
        if pot.dtype != F64 or pot.numel() != self.n:
            raise ValueError("potential must be an FP64 vector on the domain")
        self.pot = pot.contiguous()
        if self.dtype == F32:
            if self.pot32 is None:
                self.pot32 = torch.empty(self.n, dtype=F32, device="cuda")
            _lib.call("pc_convert_f64_to_f32", self.n, 1, ptr(self.pot), self.n, ptr(self.pot32), self.n, stream())

    def _v(self, dtype):
        return self.pot if dtype == F64 else self.pot32

    # ------------------------------------------------------------------ blocks
    def new_block(self, k: int, dtype=None) -> torch.Tensor:
        return torch.empty((k, self.ld), dtype=dtype or self.dtype, device="cuda")

    # ------------------------------------------------------------------ apply
    def step(self, Y: torch.Tensor, X: torch.Tensor | None, out: torch.Tensor, nvec: int,
             a: float, b: float, c: float) -> None:
        """out[:nvec] = a*H*Y + b*Y + c*X (X may be out; Y must not be)."""
        dt = Y.dtype
        if dt not in (F32, F64) or out.dtype != dt or (X is not None and X.dtype != dt):
            raise TypeError("mixed dtypes in Hamiltonian.step")
        ld = Y.shape[1]
        if out.shape[1] != ld or (X is not None and X.shape[1] != ld):
            raise ValueError("leading dimensions differ")
        if Y.shape[0] < nvec or out.shape[0] < nvec:
            raise ValueError("block narrower than nvec")
        suffix = "f64" if dt == F64 else "f32"
        V = self._v(dt)
        st = stream()
        cbuf = None
        if self.proj is not None and self.proj.natoms > 0:
            cbuf = self.scratch.get("projC", 3 * self.proj.natoms * nvec)
            _lib.call(f"pc_proj_gather_{suffix}", self.proj.handle, ptr(Y), ld, nvec, ptr(cbuf), st)
        multi = nvec >= ZMARCH_MIN_VECTORS
        kernel = f"pc_stencil_{suffix}" if multi else f"pc_stencil_direct_{suffix}"
        timed = multi and self.stencil_timer is not None
        if timed:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
        _lib.call(kernel, self.dgrid.handle, ptr(Y), ptr(X) if X is not None else C.c_void_p(0), ptr(out), ptr(V),
                  nvec, ld, a * self.kin_scale, a, b, c, C.c_void_p(0), st)
        if timed:
            e1.record()
            esz = 8 if dt == F64 else 4
            self.stencil_timer.append((e0, e1, (3 if X is not None else 2) * self.n * nvec * esz, esz))
        if cbuf is not None:
            _lib.call(f"pc_proj_scatter_{suffix}", self.proj.handle, ptr(cbuf), ptr(out), ld, nvec, a, st)

    def apply(self, Y: torch.Tensor, out: torch.Tensor, nvec: int | None = None) -> None:
        self.step(Y, None, out, Y.shape[0] if nvec is None else nvec, 1.0, 0.0, 0.0)

    def close(self):
        self.dgrid.close()


class BlockLinalg:
    """MFMA Gram / rotation GEMMs and the FP64 dense Rayleigh-Ritz stage."""

    def __init__(self, n: int, kmax: int):
        self.n = n
        self.kmax = kmax
        self.scratch = Scratch()
        handle = C.c_void_p()
        _lib.call("pc_dense_create", stream(), kmax, C.byref(handle))
        self.dense = handle

    def gram(self, X: torch.Tensor, Y: torch.Tensor, kx: int, ky: int, out: torch.Tensor, sym: bool) -> None:
        """out (ky x kx row-major == kx x ky column-major, FP64) = X^T Y."""
        nbytes = C.c_size_t()
        _lib.call("pc_gram_workspace", self.n, kx, ky, int(sym), C.byref(nbytes))
        work = self.scratch.bytes("gram", nbytes.value)
        suffix = "f64" if X.dtype == F64 else "f32"
        _lib.call(f"pc_gram_{suffix}", ptr(X), X.shape[1], ptr(Y), Y.shape[1], self.n, kx, ky, ptr(out), kx,
                  int(sym), ptr(work), nbytes.value, stream())

    def rotate(self, X: torch.Tensor, Q: torch.Tensor, kx: int, ky: int, out: torch.Tensor) -> None:
        """out[:ky] = (X[:kx]^T Q)^T, i.e. Z = X Q with Q FP64 kx x ky column-major (ld kx)."""
        nbytes = C.c_size_t()
        _lib.call("pc_rotate_workspace", kx, ky, int(X.dtype != F64), C.byref(nbytes))
        work = self.scratch.bytes("rot", nbytes.value)
        suffix = "f64" if X.dtype == F64 else "f32"
        _lib.call(f"pc_rotate_{suffix}", ptr(X), X.shape[1], ptr(Q), kx, ptr(out), out.shape[1], self.n, kx, ky,
                  ptr(work), nbytes.value, stream())

    def dense_chol_async(self, S: torch.Tensor, k: int) -> None:
        """Queue the Cholesky factor of S (k x k) on the dense stage's second
        stream; the next dense_rr with this S uses it (S must stay unchanged)."""
        _lib.call("pc_dense_chol_async", self.dense, k, ptr(S), k)

    def gram_chol_async(self, Y: torch.Tensor, k: int, S: torch.Tensor) -> None:
        """S = Y^T Y (k columns) and its Cholesky factor on the dense stage's
        second stream; the next dense_rr with this S waits for them and uses
        the factor (Y and S must stay unchanged until then)."""
        nbytes = C.c_size_t()
        _lib.call("pc_gram_workspace", self.n, k, k, 1, C.byref(nbytes))
        work = self.scratch.bytes("gram_side", nbytes.value)
        _lib.call("pc_dense_gram_chol_async", self.dense, int(Y.dtype != F64), ptr(Y), Y.shape[1], self.n, k, ptr(S),
                  k, ptr(work), nbytes.value)

    def dense_rr(self, S: torch.Tensor, G: torch.Tensor, k: int):
        """Return (T (k x k column-major FP64 device), ritz values (host numpy))."""
        import numpy as np
        T = self.scratch.get("T", k * k)
        lam_dev = self.scratch.get("lam", k)
        lam = np.zeros(k)
        _lib.call("pc_dense_rr", self.dense, k, ptr(S), k, ptr(G), k, ptr(T), k, ptr(lam_dev),
                  lam.ctypes.data_as(C.POINTER(C.c_double)))
        return T, lam

    def close(self):
        if self.dense:
            _lib.call("pc_dense_destroy", self.dense)
            self.dense = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class VecOps:
    """FP64 vector helpers (dots, linear combinations) used by Lanczos and the mixer."""

    def __init__(self, n: int):
        self.n = n
        self.scratch = Scratch()
        nch = C.c_int()
        _lib.call("pc_dots_nchunk", n, C.byref(nch))
        self.nchunk = nch.value

    def dots(self, A: torch.Tensor, m: int, v: torch.Tensor, lda: int | None = None) -> torch.Tensor:
        """Device FP64 vector out[i] = A_i . v, i < m (A rows of length >= n)."""
        lda = A.shape[-1] if lda is None else lda
        part = self.scratch.get("dots_part", m * self.nchunk)
        out = torch.empty(m, dtype=F64, device="cuda")
        _lib.call("pc_dots", self.n, ptr(A), lda, m, ptr(v), ptr(part), ptr(out), stream())
        return out

    def lincomb_dev(self, y: torch.Tensor, alpha: float = 1.0, x: torch.Tensor | None = None, beta: float = 0.0,
                    A: torch.Tensor | None = None, coef: torch.Tensor | None = None, coef_scale: float = 1.0,
                    alpha_recip: torch.Tensor | None = None, alpha_tol: float = 0.0,
                    beta_dev: torch.Tensor | None = None, beta_scale: float = 1.0) -> None:
        """lincomb with its scalars on the device (pc_dots results, no host
        round trip): coefficients coef_scale * coef, alpha = 1 / alpha_recip[0]
        (0 when |v| < alpha_tol |v|) if given, beta = beta_scale * beta_dev[0]
        if given."""
        m = 0 if coef is None else int(coef.numel())
        if m > 64:
            raise ValueError("lincomb supports at most 64 columns per call")
        nul = C.c_void_p(0)
        _lib.call("pc_lincomb_dev", self.n, ptr(A) if A is not None else nul,
                  (A.shape[-1] if A is not None else self.n), m, ptr(coef) if coef is not None else nul, coef_scale,
                  alpha, ptr(alpha_recip) if alpha_recip is not None else nul, alpha_tol,
                  ptr(x) if x is not None else nul, beta, ptr(beta_dev) if beta_dev is not None else nul,
                  beta_scale, ptr(y), stream())

    def lincomb(self, y: torch.Tensor, alpha: float = 1.0, x: torch.Tensor | None = None, beta: float = 0.0,
                A: torch.Tensor | None = None, coef=None, lda: int | None = None) -> None:
        import numpy as np
        m = 0 if coef is None else len(coef)
        cf = np.ascontiguousarray(np.asarray(coef if coef is not None else [0.0], dtype=np.float64))
        if m > 64:
            raise ValueError("lincomb supports at most 64 columns per call")
        _lib.call("pc_lincomb", self.n, ptr(A) if A is not None else C.c_void_p(0),
                  (A.shape[-1] if lda is None else lda) if A is not None else self.n, m,
                  cf.ctypes.data_as(C.POINTER(C.c_double)), alpha, ptr(x) if x is not None else C.c_void_p(0),
                  beta, ptr(y), stream())
