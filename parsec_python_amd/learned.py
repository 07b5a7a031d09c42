"""Learned SAD -> converged density initialiser (3D U-Net on the HIP kernels).

The network, its precision contract and its gradients are defined by the CPU
restatement in ``oracle/unet.py`` (4-level U-Net, 16-64 channels, leaky ReLU,
BF16 storage / FP32 accumulation).  Here every layer runs on the C-ABI kernels
``pc_unet_*``: implicit-GEMM 3x3x3 convolutions on the bf16 MFMA, pooling,
upsampling, the activation backward with the bias gradient, the weight
gradient, MSE and Adam.  Tensors are channel-last (voxel-major); the skip
tensors are written straight into the second half of the decoder's
concatenation buffers.

Input and target per volume: x = rho_SAD / s, t = (rho_conv - rho_SAD) / s with
s = max(rho_SAD); the predicted start density is rho_SAD + s * y, clipped at 0
and renormalised to the electron count.  The SCF then starts from it the way
the reference's ML-density path does (``src/V_ion/pseudoDiag_ML4Den_poisson.py:
392-540``): rho1 replaces rho0, and V_H1 = V_H0 + CG(A, 4 pi/h^3 (rho1 - rho0)).
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np
import torch

from . import _lib
from .device import F32, F64, ptr, require_gpu, stream

LEVELS = (16, 32, 48, 64)
SLOPE = 0.1
BF16 = torch.int16     # raw BF16 storage (bit patterns)


def layer_specs():
    c1, c2, c3, c4 = LEVELS
    return [("e1a", 1, c1), ("e1b", c1, c1), ("e2a", c1, c2), ("e2b", c2, c2), ("e3a", c2, c3), ("e3b", c3, c3),
            ("ba", c3, c4), ("bb", c4, c4), ("d3a", c4 + c3, c3), ("d3b", c3, c3), ("d2a", c3 + c2, c2),
            ("d2b", c2, c2), ("d1a", c2 + c1, c1), ("d1b", c1, c1), ("out", c1, 1)]


def init_params(seed: int = 0):
    """He-uniform weights W[co][ci][27], zero biases, zero output layer (start = SAD)."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, ci, co in layer_specs():
        if name == "out":
            w = np.zeros((co, ci, 27), np.float32)
        else:
            lim = math.sqrt(6.0 / (27 * ci))
            w = rng.uniform(-lim, lim, (co, ci, 27)).astype(np.float32)
        params[name] = (w, np.zeros(co, np.float32))
    return params


def _i(x):
    return int(x)


class UNet:
    """Device parameters (FP32 master + BF16 copy + Adam moments) and the
    forward / backward passes for one volume at a time."""

    def __init__(self, params=None, seed: int = 0):
        require_gpu()
        params = params if params is not None else init_params(seed)
        self.specs = layer_specs()
        self.off = {}
        tot = 0
        for name, ci, co in self.specs:
            nw = co * 27 * ci
            self.off[name] = (tot, nw, tot + nw, co)
            tot += nw + co
            tot = (tot + 7) // 8 * 8          # keep every weight block 16-byte aligned
        self.ntot = tot
        flat = np.zeros(tot, np.float32)
        for name, ci, co in self.specs:
            w, b = params[name]
            o, nw, ob, _ = self.off[name]
            flat[o:o + nw] = np.transpose(np.asarray(w, np.float32), (0, 2, 1)).reshape(-1)   # -> [co][t][ci]
            flat[ob:ob + co] = b
        dev = torch.device("cuda")
        self.w = torch.from_numpy(flat).to(dev)
        self.g = torch.zeros_like(self.w)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.wb = torch.empty(tot, dtype=BF16, device=dev)
        _lib.call("pc_unet_cast_bf16", ptr(self.w), tot, ptr(self.wb), stream())
        self.step_count = 0
        self.bufs = {}
        self.dims = None
        self.wt = {}

    # ------------------------------------------------------------------ params
    def params(self):
        flat = self.w.cpu().numpy()
        out = {}
        for name, ci, co in self.specs:
            o, nw, ob, _ = self.off[name]
            w = flat[o:o + nw].reshape(co, 27, ci).transpose(0, 2, 1).copy()
            out[name] = (w, flat[ob:ob + co].copy())
        return out

    def grads(self):
        flat = self.g.cpu().numpy()
        out = {}
        for name, ci, co in self.specs:
            o, nw, ob, _ = self.off[name]
            out[name] = (flat[o:o + nw].reshape(co, 27, ci).transpose(0, 2, 1).copy(), flat[ob:ob + co].copy())
        return out

    def save(self, path):
        p = self.params()
        np.savez(path, **{f"{k}_w": v[0] for k, v in p.items()}, **{f"{k}_b": v[1] for k, v in p.items()},
                 levels=np.array(LEVELS), slope=np.array(SLOPE))

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        if tuple(int(c) for c in z["levels"]) != LEVELS:
            raise ValueError("checkpoint has a different channel plan")
        params = {name: (z[f"{name}_w"], z[f"{name}_b"]) for name, _, _ in layer_specs()}
        return cls(params)

    def _wb(self, name):
        o, nw, _, _ = self.off[name]
        return self.wb[o:o + nw]

    def _bias(self, name):
        _, _, ob, co = self.off[name]
        return self.w[ob:ob + co]

    def _gw(self, name):
        o, nw, _, _ = self.off[name]
        return self.g[o:o + nw]

    def _gb(self, name):
        _, _, ob, co = self.off[name]
        return self.g[ob:ob + co]

    # ------------------------------------------------------------------ buffers
    def _alloc(self, D, H, W):
        if self.dims == (D, H, W):
            return
        if D % 8 or H % 8 or W % 8:
            raise ValueError("volume dimensions must be multiples of 8")
        c1, c2, c3, c4 = LEVELS
        V = [D * H * W // (8 ** l) for l in range(4)]
        dev = torch.device("cuda")

        def bf(n):
            return torch.empty(n, dtype=BF16, device=dev)

        def f32(n):
            return torch.empty(n, dtype=F32, device=dev)

        b = {}
        b["x"] = bf(V[0])
        b["e1a"] = bf(V[0] * c1)
        b["c1"] = bf(V[0] * (c2 + c1))
        b["p1"] = bf(V[1] * c1)
        b["e2a"] = bf(V[1] * c2)
        b["c2"] = bf(V[1] * (c3 + c2))
        b["p2"] = bf(V[2] * c2)
        b["e3a"] = bf(V[2] * c3)
        b["c3"] = bf(V[2] * (c4 + c3))
        b["p3"] = bf(V[3] * c3)
        b["ba"] = bf(V[3] * c4)
        b["bb"] = bf(V[3] * c4)
        b["d3a"] = bf(V[2] * c3)
        b["d3b"] = bf(V[2] * c3)
        b["d2a"] = bf(V[1] * c2)
        b["d2b"] = bf(V[1] * c2)
        b["d1a"] = bf(V[0] * c1)
        b["d1b"] = bf(V[0] * c1)
        b["y"] = f32(V[0])
        # backward
        cmax = max(ci for _, ci, _ in self.specs)
        b["gA"] = f32(V[0] * cmax)      # gradient at a conv input (FP32)
        b["gB"] = f32(V[0] * cmax)
        b["gskip1"] = f32(V[0] * (c2 + c1))
        b["gskip2"] = f32(V[1] * (c3 + c2))
        b["gskip3"] = f32(V[2] * (c4 + c3))
        b["dZ"] = bf(V[0] * max(LEVELS))
        b["dy"] = f32(V[0])
        wg = max(self._wg_bytes(co, ci) for _, ci, co in self.specs)
        b["work"] = torch.empty(max(wg, self._act_bytes(max(LEVELS), V[0]), 512 * 8), dtype=torch.uint8, device=dev)
        b["loss"] = torch.zeros(1, dtype=F64, device=dev)
        self.bufs = b
        self.dims = (D, H, W)

    @staticmethod
    def _wg_bytes(co, ci):
        n = C.c_size_t()
        _lib.call("pc_unet_wgrad_workspace", co, ci, C.byref(n))
        return n.value

    @staticmethod
    def _act_bytes(c, v):
        n = C.c_size_t()
        _lib.call("pc_unet_act_bwd_workspace", c, v, C.byref(n))
        return n.value

    # ------------------------------------------------------------------ forward
    def _conv(self, name, X, ldx, cin, Y, ldy, cout, dims, mode, yoff=0):
        D, H, W = dims
        yp = C.c_void_p(Y.data_ptr() + yoff * Y.element_size())
        _lib.call("pc_unet_conv3d_fwd", ptr(X), ldx, cin, ptr(self._wb(name)), ptr(self._bias(name)), cout, D, H, W,
                  mode, SLOPE, yp, ldy, stream())

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """x: FP32 (D, H, W) on the device -> y FP32 (D, H, W)."""
        D, H, W = x.shape
        self._alloc(D, H, W)
        b = self.bufs
        c1, c2, c3, c4 = LEVELS
        d0, d1, d2, d3 = [(D >> l, H >> l, W >> l) for l in range(4)]
        st = stream()
        _lib.call("pc_unet_cast_bf16", ptr(x.contiguous()), x.numel(), ptr(b["x"]), st)
        self._conv("e1a", b["x"], 1, 1, b["e1a"], c1, c1, d0, 1)
        self._conv("e1b", b["e1a"], c1, c1, b["c1"], c2 + c1, c1, d0, 1, yoff=c2)
        _lib.call("pc_unet_pool2_fwd", C.c_void_p(b["c1"].data_ptr() + 2 * c2), c2 + c1, c1, *d0, ptr(b["p1"]), c1, st)
        self._conv("e2a", b["p1"], c1, c1, b["e2a"], c2, c2, d1, 1)
        self._conv("e2b", b["e2a"], c2, c2, b["c2"], c3 + c2, c2, d1, 1, yoff=c3)
        _lib.call("pc_unet_pool2_fwd", C.c_void_p(b["c2"].data_ptr() + 2 * c3), c3 + c2, c2, *d1, ptr(b["p2"]), c2, st)
        self._conv("e3a", b["p2"], c2, c2, b["e3a"], c3, c3, d2, 1)
        self._conv("e3b", b["e3a"], c3, c3, b["c3"], c4 + c3, c3, d2, 1, yoff=c4)
        _lib.call("pc_unet_pool2_fwd", C.c_void_p(b["c3"].data_ptr() + 2 * c4), c4 + c3, c3, *d2, ptr(b["p3"]), c3, st)
        self._conv("ba", b["p3"], c3, c3, b["ba"], c4, c4, d3, 1)
        self._conv("bb", b["ba"], c4, c4, b["bb"], c4, c4, d3, 1)
        _lib.call("pc_unet_up2_fwd", ptr(b["bb"]), c4, c4, *d2, ptr(b["c3"]), c4 + c3, st)
        self._conv("d3a", b["c3"], c4 + c3, c4 + c3, b["d3a"], c3, c3, d2, 1)
        self._conv("d3b", b["d3a"], c3, c3, b["d3b"], c3, c3, d2, 1)
        _lib.call("pc_unet_up2_fwd", ptr(b["d3b"]), c3, c3, *d1, ptr(b["c2"]), c3 + c2, st)
        self._conv("d2a", b["c2"], c3 + c2, c3 + c2, b["d2a"], c2, c2, d1, 1)
        self._conv("d2b", b["d2a"], c2, c2, b["d2b"], c2, c2, d1, 1)
        _lib.call("pc_unet_up2_fwd", ptr(b["d2b"]), c2, c2, *d0, ptr(b["c1"]), c2 + c1, st)
        self._conv("d1a", b["c1"], c2 + c1, c2 + c1, b["d1a"], c1, c1, d0, 1)
        self._conv("d1b", b["d1a"], c1, c1, b["d1b"], c1, c1, d0, 1)
        self._conv("out", b["d1b"], c1, c1, b["y"], 1, 1, d0, 0)
        return b["y"].view(D, H, W)

    # ------------------------------------------------------------------ backward
    def _wt(self, name, ci, co):
        key = name
        if key not in self.wt or self.wt[key].numel() != co * 27 * ci:
            self.wt[key] = torch.empty(co * 27 * ci, dtype=BF16, device="cuda")
        _lib.call("pc_unet_conv3d_wflip", ptr(self._wb(name)), co, ci, ptr(self.wt[key]), stream())
        return self.wt[key]

    def _layer_bwd(self, name, gA, ldg, a_out, lda_out, ci, co, X, ldx, dims, gX=None, need_dx=True):
        """From the gradient at this layer's (activated) output, produce the
        weight/bias gradients and (optionally) the FP32 gradient at its input."""
        D, H, W = dims
        b = self.bufs
        st = stream()
        work = b["work"]
        _lib.call("pc_unet_act_bwd", ptr(gA), ldg, a_out, lda_out, co, D, H, W, SLOPE, ptr(b["dZ"]), co,
                  ptr(self._gb(name)), ptr(work), work.numel(), st)
        _lib.call("pc_unet_conv3d_wgrad", ptr(b["dZ"]), co, co, X, ldx, ci, D, H, W, ptr(self._gw(name)),
                  ptr(work), work.numel(), st)
        if need_dx:
            wt = self._wt(name, ci, co)
            _lib.call("pc_unet_conv3d_fwd", ptr(b["dZ"]), co, co, ptr(wt), None, ci, D, H, W, 2, SLOPE, ptr(gX), ci,
                      st)

    def backward(self, dy: torch.Tensor) -> None:
        """dy: FP32 (D, H, W) gradient of the loss w.r.t. y; fills self.g."""
        D, H, W = self.dims
        b = self.bufs
        c1, c2, c3, c4 = LEVELS
        d0, d1, d2, d3 = [(D >> l, H >> l, W >> l) for l in range(4)]
        st = stream()
        gA, gB = b["gA"], b["gB"]

        def at(t, ch):   # pointer to channel offset ch of a BF16 buffer
            return C.c_void_p(t.data_ptr() + 2 * ch)

        def atf(t, ch):
            return C.c_void_p(t.data_ptr() + 4 * ch)

        # out (no activation)
        self._layer_bwd("out", dy.contiguous(), 1, None, 0, c1, 1, ptr(b["d1b"]), c1, d0, gA)
        self._layer_bwd("d1b", gA, c1, ptr(b["d1b"]), c1, c1, c1, ptr(b["d1a"]), c1, d0, gB)
        self._layer_bwd("d1a", gB, c1, ptr(b["d1a"]), c1, c2 + c1, c1, ptr(b["c1"]), c2 + c1, d0, b["gskip1"])
        _lib.call("pc_unet_up2_bwd", ptr(b["gskip1"]), c2 + c1, c2, *d0, ptr(gA), c2, st)
        self._layer_bwd("d2b", gA, c2, ptr(b["d2b"]), c2, c2, c2, ptr(b["d2a"]), c2, d1, gB)
        self._layer_bwd("d2a", gB, c2, ptr(b["d2a"]), c2, c3 + c2, c2, ptr(b["c2"]), c3 + c2, d1, b["gskip2"])
        _lib.call("pc_unet_up2_bwd", ptr(b["gskip2"]), c3 + c2, c3, *d1, ptr(gA), c3, st)
        self._layer_bwd("d3b", gA, c3, ptr(b["d3b"]), c3, c3, c3, ptr(b["d3a"]), c3, d2, gB)
        self._layer_bwd("d3a", gB, c3, ptr(b["d3a"]), c3, c4 + c3, c3, ptr(b["c3"]), c4 + c3, d2, b["gskip3"])
        _lib.call("pc_unet_up2_bwd", ptr(b["gskip3"]), c4 + c3, c4, *d2, ptr(gA), c4, st)
        self._layer_bwd("bb", gA, c4, ptr(b["bb"]), c4, c4, c4, ptr(b["ba"]), c4, d3, gB)
        self._layer_bwd("ba", gB, c4, ptr(b["ba"]), c4, c3, c4, ptr(b["p3"]), c3, d3, gA)
        # e3 output gradient = pool backward + skip gradient
        _lib.call("pc_unet_pool2_bwd", ptr(gA), c3, atf(b["gskip3"], c4), c4 + c3, c3, *d2, ptr(gB), c3, st)
        self._layer_bwd("e3b", gB, c3, at(b["c3"], c4), c4 + c3, c3, c3, ptr(b["e3a"]), c3, d2, gA)
        self._layer_bwd("e3a", gA, c3, ptr(b["e3a"]), c3, c2, c3, ptr(b["p2"]), c2, d2, gB)
        _lib.call("pc_unet_pool2_bwd", ptr(gB), c2, atf(b["gskip2"], c3), c3 + c2, c2, *d1, ptr(gA), c2, st)
        self._layer_bwd("e2b", gA, c2, at(b["c2"], c3), c3 + c2, c2, c2, ptr(b["e2a"]), c2, d1, gB)
        self._layer_bwd("e2a", gB, c2, ptr(b["e2a"]), c2, c1, c2, ptr(b["p1"]), c1, d1, gA)
        _lib.call("pc_unet_pool2_bwd", ptr(gA), c1, atf(b["gskip1"], c2), c2 + c1, c1, *d0, ptr(gB), c1, st)
        self._layer_bwd("e1b", gB, c1, at(b["c1"], c2), c2 + c1, c1, c1, ptr(b["e1a"]), c1, d0, gA)
        self._layer_bwd("e1a", gA, c1, ptr(b["e1a"]), c1, 1, c1, ptr(b["x"]), 1, d0, None, need_dx=False)

    # ------------------------------------------------------------------ training
    def loss_grad(self, y: torch.Tensor, t: torch.Tensor) -> torch.Tensor:
        """MSE loss (device FP64 scalar tensor) and its gradient into bufs['dy']."""
        b = self.bufs
        n = y.numel()
        _lib.call("pc_unet_mse", ptr(y), ptr(t.contiguous()), n, ptr(b["dy"]), ptr(b["loss"]), ptr(b["work"]),
                  b["work"].numel(), stream())
        return b["loss"][0] / n

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.step_count += 1
        _lib.call("pc_unet_adam", ptr(self.w), ptr(self.g), ptr(self.m), ptr(self.v), self.ntot, lr, b1, b2, eps, wd,
                  self.step_count, ptr(self.wb), stream())

    def train_step(self, x, t, lr=1e-3):
        y = self.forward(x)
        loss = self.loss_grad(y, t)
        self.backward(self.bufs["dy"].view_as(y))
        self.adam_step(lr)
        return loss


# ---------------------------------------------------------------------- use in the SCF
def pad_to(a: torch.Tensor, mult: int = 8):
    D, H, W = a.shape
    P = [(-s) % mult for s in (D, H, W)]
    if not any(P):
        return a, (D, H, W)
    out = torch.zeros((D + P[0], H + P[1], W + P[2]), dtype=a.dtype, device=a.device)
    out[:D, :H, :W] = a
    return out, (D, H, W)


def features(grid, rho0_cell: torch.Tensor):
    """SAD density (electrons per cell, domain vector) -> (x volume, scale)."""
    h3 = grid.h ** 3
    full = torch.zeros(grid.nx * grid.ny * grid.nz, dtype=F64, device="cuda")
    idx = torch.as_tensor(grid.flat_index, dtype=torch.long, device="cuda")
    full[idx] = rho0_cell / h3
    s = float(full.max().item())
    vol = (full / s).to(F32).view(grid.nz, grid.ny, grid.nx)
    return vol, s, idx


def predict_start(net: UNet, grid, rho0_cell: torch.Tensor, nelec: float) -> torch.Tensor:
    """Learned start density (electrons per cell on the domain), normalised to nelec."""
    x, s, idx = features(grid, rho0_cell)
    xp, (D, H, W) = pad_to(x)
    y = net.forward(xp)[:D, :H, :W].reshape(-1).to(F64)
    rho = rho0_cell / (grid.h ** 3) + s * y[idx]
    rho = torch.clamp(rho, min=0.0) * grid.h ** 3
    return rho * (nelec / float(rho.sum().item()))


def learned_start(prob, net: UNet, cg_maxit: int | None = None, cg_tol: float = 1e-5):
    """(rho1, V_H1) for run_scf(rho_init=..., hpot_init=...), following the ML
    start of src/V_ion/pseudoDiag_ML4Den_poisson.py:392-540: rho1 is the
    clipped, renormalised network density; V_H1 = V_H0 + CG(A, 4 pi/h^3 (rho1 - rho0))
    from a zero start to the reference's relative tolerance 1e-5 (:532-538); the
    iteration cap defaults to the problem's Hartree cap."""
    if cg_maxit is None:
        cg_maxit = int(getattr(prob.cfg, "pcg_maxit", 200))
    from .hartree import HartreeSolver
    rho1 = predict_start(net, prob.grid, prob.rho0, prob.nelec)
    h3 = prob.h ** 3
    rhs = (4.0 * math.pi / h3) * (rho1 - prob.rho0)
    dv = torch.zeros_like(rhs)
    hs = HartreeSolver(prob.H.dgrid, prob.h, reference_aliasing=False, prec=getattr(prob.cfg, "pcg_prec", "none"))
    hs.solve(rhs, dv, cg_maxit, cg_tol)
    hs.close()
    return rho1, prob.hpot0 + dv


def default_checkpoint() -> str:
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "unet_density.npz")


# ---------------------------------------------------------------------- training data (own SCF runs)
TRAIN_H_ANG = 0.25        # grid spacing of the training volumes (the Si87H76 spacing)
TRAIN_RADIUS_ANG = 7.9    # -> nx = 64 cube


def distorted_si_cluster(n_si: int, rng: np.random.Generator, jitter_ang=0.08, scale_jitter=0.02):
    from . import molecules as M
    a = 5.43 * (1.0 + rng.uniform(-scale_jitter, scale_jitter))
    geo = M.silicon_cluster(a=a, n_si=n_si)
    return [(s, x + rng.uniform(-jitter_ang, jitter_ang, x.shape)) for s, x in geo]


def sample_geometry(kind: str, rng: np.random.Generator):
    from . import molecules as M
    if kind == "alkane":
        return M.alkane(int(rng.integers(1, 9)), rng)
    if kind == "si":
        return distorted_si_cluster(int(rng.choice([5, 17, 35])), rng)
    raise ValueError(kind)


def scf_pair(geo, precision="mixed", etol=1e-6, domain="cube", max_wall=None):
    """(grid, rho_SAD per cell, converged rho per cell, result, problem) from one SCF run."""
    from . import molecules as M
    from .scf import SCFConfig, SCFProblem, run_scf
    atoms, dom, nev = M.prepare(geo, h_ang=TRAIN_H_ANG, radius_ang=TRAIN_RADIUS_ANG)
    cfg = SCFConfig(tol=1e-12, maxits=60, domain=domain, precision=precision, etol_ha_per_atom=etol,
                    pcg_maxit=1000, pcg_aliasing=False, seed=1)
    prob = SCFProblem(atoms, dom, nev, cfg)
    res = run_scf(prob, max_wall=max_wall)
    return prob, res


def volume_pair(prob, res):
    """Network input / target volumes (FP32, nz x ny x nx) and the scale s."""
    x, s, idx = features(prob.grid, prob.rho0)
    full = torch.zeros(prob.grid.nx * prob.grid.ny * prob.grid.nz, dtype=F64, device="cuda")
    full[idx] = res.rho - prob.rho0 / prob.h ** 3
    t = (full / s).to(F32).view_as(x)
    return x, t, s


def iterations_with(geo, net: UNet | None, precision="mixed", etol=1e-6, domain="cube", seed=1,
                    sad_weighted=False):
    """SCF iterations and energy from the SAD start (net None; with
    ``sad_weighted`` the first_filt start vectors are weighted by the SAD
    density, the control for the learned start's weighting) or the learned
    start."""
    from . import molecules as M
    from .scf import SCFConfig, SCFProblem, run_scf
    atoms, dom, nev = M.prepare(geo, h_ang=TRAIN_H_ANG, radius_ang=TRAIN_RADIUS_ANG)
    cfg = SCFConfig(tol=1e-12, maxits=60, domain=domain, precision=precision, etol_ha_per_atom=etol,
                    pcg_maxit=1000, pcg_aliasing=False, seed=seed)
    prob = SCFProblem(atoms, dom, nev, cfg)
    if net is None:
        res = run_scf(prob, start_weight_rho=prob.rho0 if sad_weighted else None)
    else:
        rho1, vh1 = learned_start(prob, net)
        res = run_scf(prob, rho_init=rho1, hpot_init=vh1)
    out = (res.iterations, 0.5 * res.energies["total_ry"] / prob.n_atoms, res.converged)
    prob.close()
    return out


def evaluate_main(argv=None):
    """Held-out evaluation of a checkpoint: for fresh seeded geometries (the
    training kinds, seeds it never saw) the start density's L1 distance to the
    converged one and the mixed-precision SCF iterations to |dE| < 1e-6
    Ha/atom from the SAD start, the SAD start with SAD-weighted start vectors
    (the control) and the learned start (weighted by the learned density),
    each for several start-vector seeds; writes a Markdown report."""
    import argparse
    import time
    from . import molecules as M
    from .scf import SCFConfig, SCFProblem, run_scf
    ap = argparse.ArgumentParser(prog="python -m parsec_python_amd.learned evaluate")
    ap.add_argument("--checkpoint", default=None)
    ap.add_argument("--geometries", type=int, default=8)
    ap.add_argument("--seeds", type=int, default=3)
    ap.add_argument("--heldout-seed", type=int, default=777001)
    ap.add_argument("--out", default="unet_heldout_report.md")
    a = ap.parse_args(argv)
    net = UNet.load(a.checkpoint or default_checkpoint())
    rng = np.random.default_rng(a.heldout_seed)
    rows = []
    t0 = time.perf_counter()
    for g in range(a.geometries):
        kind = "si" if g % 4 == 3 else "alkane"
        geo = sample_geometry(kind, rng)
        atoms, dom, nev = M.prepare(geo, h_ang=TRAIN_H_ANG, radius_ang=TRAIN_RADIUS_ANG)
        its = {"sad": [], "sad_weighted": [], "learned": []}
        e = {"sad": [], "sad_weighted": [], "learned": []}
        l1 = None
        for sd in range(1, a.seeds + 1):
            cfg = SCFConfig(tol=1e-12, maxits=60, domain="cube", precision="mixed", etol_ha_per_atom=1e-6,
                            pcg_maxit=1000, pcg_aliasing=False, seed=sd)
            prob = SCFProblem(atoms, dom, nev, cfg)
            r0 = run_scf(prob)
            r0w = run_scf(prob, start_weight_rho=prob.rho0)
            rho1, vh1 = learned_start(prob, net)
            r1 = run_scf(prob, rho_init=rho1, hpot_init=vh1)
            if l1 is None:
                rc = r0.rho * prob.h ** 3
                l1 = (float((prob.rho0 - rc).abs().sum()) / prob.nelec, float((rho1 - rc).abs().sum()) / prob.nelec)
            for key, r in (("sad", r0), ("sad_weighted", r0w), ("learned", r1)):
                its[key].append(r.iterations)
                e[key].append(0.5 * r.energies["total_ry"] / prob.n_atoms)
            natoms = prob.n_atoms
            prob.close()
        de = max(abs(x - e["sad"][0]) for k in e for x in e[k])
        rows.append({"kind": kind, "atoms": natoms, "points": dom["nx"] ** 3, "l1": l1, "its": its, "de": de})
        print(f"[learned] held-out {g}: {kind} ({natoms} atoms) L1 SAD {l1[0]:.3f} learned {l1[1]:.4f}; its "
              f"SAD {its['sad']} SAD-weighted {its['sad_weighted']} learned {its['learned']}", flush=True)
    mean = {k: float(np.mean([np.mean(r["its"][k]) for r in rows])) for k in ("sad", "sad_weighted", "learned")}
    lines = ["# Learned start: held-out evaluation", "",
             f"Checkpoint `{os.path.basename(a.checkpoint or default_checkpoint())}`; {a.geometries} geometries drawn "
             f"with seed {a.heldout_seed} (distorted alkanes CnH2n+2 and H-passivated Si clusters, the training kinds, "
             f"seeds the training never used), cube {TRAIN_RADIUS_ANG} A radius at h = {TRAIN_H_ANG} A (64^3); "
             f"mixed precision, stop at |dE| < 1e-6 Ha/atom; start-vector seeds 1..{a.seeds}. "
             f"Command: `python -m parsec_python_amd.learned evaluate --geometries {a.geometries} --seeds {a.seeds}`. "
             f"Wall time {time.perf_counter() - t0:.0f} s.", "",
             "| # | kind | atoms | start L1 / e- SAD | learned | SCF its SAD | SAD + SAD-weighted vectors | learned "
             "(+ learned-weighted vectors) | max energy spread (Ha/atom) |",
             "|---|---|---|---|---|---|---|---|---|"]
    for i, r in enumerate(rows):
        lines.append(f"| {i} | {r['kind']} | {r['atoms']} | {r['l1'][0]:.3f} | {r['l1'][1]:.4f} | "
                     f"{' '.join(map(str, r['its']['sad']))} | {' '.join(map(str, r['its']['sad_weighted']))} | "
                     f"{' '.join(map(str, r['its']['learned']))} | {r['de']:.1e} |")
    lines += ["", f"Mean SCF iterations: SAD {mean['sad']:.2f}, SAD with SAD-weighted start vectors "
                  f"{mean['sad_weighted']:.2f}, learned start {mean['learned']:.2f}.", ""]
    with open(a.out, "w") as f:
        f.write("\n".join(lines))
    print(mean, flush=True)
    return rows, mean


def train_main(argv=None):
    import argparse
    import json
    import time
    ap = argparse.ArgumentParser(prog="python -m parsec_python_amd.learned")
    ap.add_argument("--alkanes", type=int, default=400)
    ap.add_argument("--si", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--gen-budget-s", type=float, default=480.0, help="stop generating data after this long")
    ap.add_argument("--train-budget-s", type=float, default=300.0)
    ap.add_argument("--heldout", type=int, default=8)
    ap.add_argument("--out", default="unet_density.npz")
    ap.add_argument("--report", default="unet_training.json")
    a = ap.parse_args(argv)
    rng = np.random.default_rng(a.seed)
    kinds = ["alkane"] * a.alkanes + ["si"] * a.si
    rng.shuffle(kinds)
    xs, ts, meta = [], [], []
    t0 = time.perf_counter()
    for i, kind in enumerate(kinds):
        if time.perf_counter() - t0 > a.gen_budget_s:
            break
        geo = sample_geometry(kind, rng)
        prob, res = scf_pair(geo)
        x, t, s = volume_pair(prob, res)
        xs.append(x.to(torch.bfloat16))     # storage only; the network input is re-cast from FP32 below
        ts.append(t.to(torch.bfloat16))
        meta.append({"kind": kind, "atoms": int(prob.n_atoms), "its": res.iterations, "converged": res.converged})
        prob.close()
        if (i + 1) % 20 == 0:
            print(f"[learned] data {i + 1}/{len(kinds)}  {time.perf_counter() - t0:.0f} s", flush=True)
    t_gen = time.perf_counter() - t0
    n = len(xs)
    print(f"[learned] {n} samples in {t_gen:.0f} s", flush=True)
    net = UNet(seed=a.seed)
    order = np.arange(n)
    losses = []
    t1 = time.perf_counter()
    ep = 0
    for ep in range(a.epochs):
        rng.shuffle(order)
        lr = a.lr * (0.5 * (1 + math.cos(math.pi * ep / a.epochs)))
        tot = torch.zeros((), dtype=F64, device="cuda")
        for j in order:
            tot += net.train_step(xs[j].float(), ts[j].float(), lr=lr)
        losses.append(float(tot.item()) / n)
        print(f"[learned] epoch {ep + 1}: loss {losses[-1]:.3e}  {time.perf_counter() - t1:.0f} s", flush=True)
        if time.perf_counter() - t1 > a.train_budget_s:
            break
    t_train = time.perf_counter() - t1
    net.save(a.out)
    # held-out geometries (fresh seed): SCF iterations from SAD vs learned start
    hrng = np.random.default_rng(a.seed + 1)
    held = []
    for k in range(a.heldout):
        kind = "si" if k % 4 == 3 else "alkane"
        geo = sample_geometry(kind, hrng)
        its0, e0, c0 = iterations_with(geo, None)
        its1, e1, c1 = iterations_with(geo, net)
        held.append({"kind": kind, "sad_its": its0, "learned_its": its1, "dE_ha_atom": abs(e1 - e0),
                     "converged": bool(c0 and c1)})
        print(f"[learned] held-out {k}: {kind} SAD {its0} its, learned {its1} its, |dE| {abs(e1 - e0):.2e}",
              flush=True)
    rep = {"samples": n, "gen_s": t_gen, "train_s": t_train, "epochs": ep + 1, "loss": losses,
           "heldout": held, "data": meta}
    with open(a.report, "w") as f:
        json.dump(rep, f)
    print(json.dumps({k: rep[k] for k in ("samples", "gen_s", "train_s", "epochs")}), flush=True)


if __name__ == "__main__":
    import sys
    if len(sys.argv) > 1 and sys.argv[1] == "evaluate":
        evaluate_main(sys.argv[2:])
    else:
        train_main()
