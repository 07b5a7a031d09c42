"""Kernel workloads for timing and hardware-counter runs (one GPU).

    python -m parsec_python_amd.kernel_bench synthetic [--dtype f32] [--reps R]
        The synthetic kernel workload: a 128^3 cube restricted to the index
        sphere (i-63.5)^2+(j-63.5)^2+(k-63.5)^2 <= 60^2 (904,960 points),
        h = 0.377945 bohr, fd order 12 (radius 6), V_loc = 20 Gaussians
        -2 exp(-|r-R_a|^2 / 2) Ha (width 1 bohr) with R_a uniform in the ball of
        radius 40 points (numpy PCG64 seed 20261202), 256 random-normal vectors
        (counter generator, seed 20261202); bounds from 6 Lanczos steps, a
        degree-16 scaled Chebyshev filter, then the Rayleigh-Ritz Gram /
        projected-Gram / rotation GEMMs at 904,960 x 256.  Prints one JSON line.
    python -m parsec_python_amd.kernel_bench stencil --workload si87h76 [--dtype f32] [--reps R]
        R three-term recurrence steps Y_{j+1} = a H Y_j + b Y_j + c Y_{j-1} on the
        workload's sphere domain with its block width (no projectors).
    python -m parsec_python_amd.kernel_bench cg --workload si87h76 [--reps R]
        300 Hartree CG iterations (FP64) on the workload's sphere domain.
    python -m parsec_python_amd.kernel_bench copy [--mb 666] [--reps R]
        A plain device-to-device copy of known size, the calibration run for
        the FETCH_SIZE / WRITE_SIZE counters.
"""
from __future__ import annotations

import argparse
import json

import numpy as np
import torch

from .device import F32, F64, require_gpu
from .grid import Grid, make_grid, make_index_sphere_grid

SYNTH_SEED = 20261202
SYNTH_H = 0.377945


def synthetic_grid(n: int = 128, radius_pts: float = 60.0) -> Grid:
    c = 0.5 * (n - 1)
    return make_index_sphere_grid(n, c, radius_pts, SYNTH_H)


def synthetic_potential(grid: Grid, n_gauss: int = 20, ball_pts: float = 40.0, seed: int = SYNTH_SEED) -> np.ndarray:
    """V_loc(r) = sum_a -2 exp(-|r - R_a|^2 / 2) Ha on the domain points (bohr)."""
    rng = np.random.default_rng(seed)
    c = 0.5 * (grid.nx - 1)
    centres = []
    while len(centres) < n_gauss:           # uniform in the ball by rejection
        p = rng.uniform(-ball_pts, ball_pts, 3)
        if p @ p <= ball_pts * ball_pts:
            centres.append(p)
    centres = (np.array(centres) + c) * grid.h
    flat = grid.flat_index
    i = flat % grid.nx
    j = (flat // grid.nx) % grid.ny
    k = flat // (grid.nx * grid.ny)
    pts = np.stack([i, j, k], axis=1).astype(np.float64) * grid.h
    v = np.zeros(grid.npts)
    for R in centres:
        d2 = np.sum((pts - R) ** 2, axis=1)
        v += -2.0 * np.exp(-0.5 * d2)
    return v


def workload_grid(name: str):
    """(grid, block width) of a named SCF workload's sphere domain."""
    from . import molecules as M
    table = {"si87h76": (M.silicon_cluster, 0.25, 13.0, 260), "c10h8": (M.naphthalene, 0.2, 9.0, 34),
             "h2o": (M.h2o, 0.2, 5.0, 8)}
    geo, h, r, nev = table[name]
    _, dom, nev = M.prepare(geo(), h_ang=h, radius_ang=r, nev=nev)
    return make_grid(int(dom["nx"]), float(dom["h"]), float(dom["radius"]), "sphere"), nev + 15


def _events():
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def run_synthetic(dtype=F32, reps: int = 1, n: int = 128, radius_pts: float = 60.0, nvec: int = 256,
                  degree: int = 16, fd_radius: int = 6) -> dict:
    from .eigensolvers import Eigensolver, rand_block
    from .hamiltonian import Hamiltonian
    require_gpu()
    g = synthetic_grid(n, radius_pts)
    pot = torch.from_numpy(synthetic_potential(g)).to("cuda")
    H = Hamiltonian(g, fd_radius, None, pot, dtype)
    es = Eigensolver(H, nvec, seed=SYNTH_SEED)
    upper, ritz = es.lancz_uppbnd(6)
    lam1 = float(ritz[0])
    low = 0.8 * float(ritz[0]) + 0.2 * float(ritz[-2])
    X = es.bufs[0]
    out = {"points": g.npts, "nvec": nvec, "degree": degree, "fd_radius": fd_radius,
           "bounds": [lam1, low, upper]}
    s = 4 if dtype == F32 else 8
    times = {"filter": [], "gram_sym": [], "gram": [], "dense_rr": [], "rotate": []}
    for _ in range(reps):
        rand_block(X, g.npts, nvec, SYNTH_SEED, normal=True)
        ev = [_events() for _ in range(5)]
        ev[0][0].record()
        Y = es.ch_filter(X, nvec, degree, lam1, low, upper)
        ev[0][1].record()
        HY = es._other(Y)
        H.apply(Y, HY, nvec)
        S = es.S.view(-1)[:nvec * nvec]
        G = es.G.view(-1)[:nvec * nvec]
        ev[1][0].record()
        es.la.gram(Y, Y, nvec, nvec, S, sym=True)
        ev[1][1].record()
        ev[2][0].record()
        es.la.gram(Y, HY, nvec, nvec, G, sym=False)
        ev[2][1].record()
        # the dense stage with rocSOLVER's Jacobi (method 1) and divide and
        # conquer (method 0) eigensolvers, on copies (timed separately), then
        # the default (method 2: one-launch tridiagonalization, this library's
        # tridiagonal eigensolver and back-transformation), which the SCF runs
        from . import _lib
        lam_alt = {}
        for method, key in ((1, "dense_rr_jacobi"), (0, "dense_rr_syevd")):
            S2, G2 = S.clone(), G.clone()
            _lib.call("pc_dense_set_method", es.la.dense, method)
            e0, e1 = _events()
            e0.record()
            _, lam_alt[key] = es.la.dense_rr(S2, G2, nvec)
            e1.record()
            torch.cuda.synchronize()
            times.setdefault(key, []).append(e0.elapsed_time(e1))
        _lib.call("pc_dense_set_method", es.la.dense, 2)
        ev[3][0].record()
        T, lam = es.la.dense_rr(S, G, nvec)
        ev[3][1].record()
        torch.cuda.synchronize()
        for key, lam_k in lam_alt.items():
            out[f"{key}_max_ritz_diff"] = float(np.max(np.abs(np.asarray(lam_k) - np.asarray(lam))))
        ev[4][0].record()
        es.la.rotate(Y, T, nvec, nvec, HY)
        ev[4][1].record()
        torch.cuda.synchronize()
        for key, (a, b) in zip(["filter", "gram_sym", "gram", "dense_rr", "rotate"], ev):
            times[key].append(a.elapsed_time(b))
    ms = {k: float(np.min(v)) for k, v in times.items()}
    fbytes = g.npts * nvec * s * (2 + 3 * (degree - 1))
    flops = 2.0 * g.npts * nvec * nvec
    out.update({
        "ms": ms,
        "filter_GBs": fbytes / (ms["filter"] * 1e-3) / 1e9,
        "filter_bytes": fbytes,
        "gram_TFs": flops / (ms["gram"] * 1e-3) / 1e12,
        "gram_sym_TFs": flops / (ms["gram_sym"] * 1e-3) / 1e12,
        "rotate_TFs": flops / (ms["rotate"] * 1e-3) / 1e12,
        "ritz_lowest": [float(x) for x in lam[:4]],
    })
    es.close()
    H.close()
    return out


def run_stencil(name: str, dtype=F32, reps: int = 10) -> dict:
    from .hamiltonian import Hamiltonian
    require_gpu()
    g, k = workload_grid(name)
    H = Hamiltonian(g, 4, None, torch.zeros(g.npts, dtype=F64, device="cuda"), dtype)
    Y = H.new_block(k)
    X = H.new_block(k)
    Y.normal_()
    X.normal_()
    H.step(Y, X, X, k, 0.3, 0.1, -1.0)
    torch.cuda.synchronize()
    e0, e1 = _events()
    e0.record()
    for _ in range(reps):
        H.step(Y, X, X, k, 0.3, 0.1, -1.0)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = 3 * g.npts * k * (4 if dtype == F32 else 8)
    H.close()
    return {"workload": name, "points": g.npts, "nvec": k, "dtype": str(dtype)[6:], "ms": ms,
            "bytes": nbytes, "GBs": nbytes / (ms * 1e-3) / 1e9}


def run_cg(name: str, iters: int = 300, reps: int = 3, prec: str | None = None, tol: float = 0.0) -> dict:
    """Hartree CG on the workload's domain: rhs ~ N(0,1), zero start, fixed
    iteration count (tolerance 0), so the time per iteration is measured."""
    from .grid import DeviceGrid
    from .hartree import HartreeSolver
    require_gpu()
    g, _ = workload_grid(name)
    dg = DeviceGrid(g, 4)
    hs = HartreeSolver(dg, g.h, reference_aliasing=prec is None, check_every=16, prec=prec)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(SYNTH_SEED)
    rhs = torch.randn(g.npts, dtype=F64, device="cuda", generator=gen)
    best = 1e30
    its = 0
    for _ in range(reps):
        x = torch.zeros(g.npts, dtype=F64, device="cuda")
        torch.cuda.synchronize()
        e0, e1 = _events()
        e0.record()
        its = hs.solve(rhs, x, iters, tol)
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    hs.close()
    dg.close()
    return {"workload": name, "points": g.npts, "prec": prec or "none", "tol": tol, "iterations": its, "ms": best,
            "us_per_iteration": 1e3 * best / max(its, 1), "final_rel_residual": hs.last_rel}


def run_copy(mb: float = 666.0, reps: int = 3) -> dict:
    require_gpu()
    n = int(mb * 1e6) // 4
    a = torch.randn(n, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    e0, e1 = _events()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return {"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "ms": ms, "GBs": 8 * n / (ms * 1e-3) / 1e9}


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m parsec_python_amd.kernel_bench")
    ap.add_argument("what", choices=["synthetic", "stencil", "cg", "copy"])
    ap.add_argument("--workload", default="si87h76", choices=["si87h76", "c10h8", "h2o"])
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"])
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--mb", type=float, default=666.0)
    ap.add_argument("--prec", default="none", choices=["none", "mg"], help="cg: preconditioner")
    ap.add_argument("--tol", type=float, default=0.0, help="cg: relative tolerance (0: run --iters iterations)")
    ap.add_argument("--iters", type=int, default=300, help="cg: iteration cap")
    a = ap.parse_args(argv)
    dt = F32 if a.dtype == "f32" else F64
    if a.what == "synthetic":
        res = run_synthetic(dt, a.reps)
    elif a.what == "stencil":
        res = run_stencil(a.workload, dt, a.reps)
    elif a.what == "cg":
        res = run_cg(a.workload, a.iters, a.reps, None if a.prec == "none" else a.prec, a.tol)
    else:
        res = run_copy(a.mb, a.reps)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
