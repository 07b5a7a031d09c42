"""GPU parity tests: each HIP entry point against the FP64 oracle (or a closed
form) on seeded inputs.  Tolerances: FP64 paths 1e-11 relative (reduction order
only), FP32 paths 2e-5 relative to the operand scale."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fd as ofd
from oracle import grid as ogrid
from oracle import pseudo as opseudo
from oracle import xc as oxc
from oracle.pcg import pcg as opcg

pytestmark = pytest.mark.gpu


def _grid(nx, ny, nz, h, kind):
    from parsec_python_amd.grid import Grid, make_grid
    if kind == "cube" and not (nx == ny == nz):
        return Grid(nx=nx, ny=ny, nz=nz, h=h, rad=0.5 * h * (nx - 1), mask=np.ones(nx * ny * nz, bool), kind="cube")
    return make_grid(nx, h, 0.5 * h * (nx - 1), kind)


def _oracle_lap(g, radius):
    return ofd.laplacian_masked(g.nx, g.ny, g.nz, radius, g.mask)


CASES = [  # nx, ny, nz, kind, radius
    (20, 20, 20, "cube", 4), (21, 21, 21, "sphere", 4), (18, 18, 18, "sphere", 6), (16, 16, 16, "cube", 1),
    (17, 17, 17, "cube", 2), (19, 19, 19, "sphere", 3), (70, 12, 9, "cube", 4), (130, 9, 7, "cube", 6),
    # rows of 65..128 points (the FP32 paired kernel): chords starting at odd
    # and even x, first and last points of the array at both parities
    (67, 67, 67, "sphere", 3), (75, 75, 75, "sphere", 1), (66, 66, 66, "sphere", 6), (97, 11, 13, "cube", 2),
    (106, 106, 106, "sphere", 4),
]


@pytest.mark.parametrize("nx,ny,nz,kind,radius", CASES)
@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("nvec", [1, 2, 9])
def test_stencil_step_vs_oracle(gpu, nx, ny, nz, kind, radius, dtype, nvec):
    import torch
    from parsec_python_amd.hamiltonian import Hamiltonian, leading_dim
    print(f"stencil case {nx}x{ny}x{nz} {kind} r={radius} {dtype} nvec={nvec}", flush=True)
    h = 0.45
    g = _grid(nx, ny, nz, h, kind)
    n = g.npts
    rng = np.random.default_rng(nx * 31 + radius)
    V = rng.standard_normal(n)
    Y = rng.standard_normal((nvec, n))
    X = rng.standard_normal((nvec, n))
    a, b, c = 0.7, -0.3, -1.1
    tdt = torch.float64 if dtype == "f64" else torch.float32
    H = Hamiltonian(g, radius, None, torch.tensor(V, device="cuda"), tdt)
    ld = leading_dim(n)
    Yd = torch.zeros((nvec, ld), dtype=tdt, device="cuda")
    Xd = torch.zeros((nvec, ld), dtype=tdt, device="cuda")
    Yd[:, :n] = torch.tensor(Y, dtype=tdt)
    Xd[:, :n] = torch.tensor(X, dtype=tdt)
    # in-place form: out aliases X
    H.step(Yd, Xd, Xd, nvec, a, b, c)
    got = Xd[:, :n].double().cpu().numpy()
    L = _oracle_lap(g, radius) * (0.5 / h ** 2)
    Yr = Y.astype(np.float32).astype(np.float64) if dtype == "f32" else Y
    Xr = X.astype(np.float32).astype(np.float64) if dtype == "f32" else X
    Vr = V.astype(np.float32).astype(np.float64) if dtype == "f32" else V
    ref = a * (L @ Yr.T).T + (a * Vr + b) * Yr + c * Xr
    tol = 1e-11 if dtype == "f64" else 2e-5
    scale = np.abs(ref).max()
    assert np.max(np.abs(got - ref)) <= tol * scale, np.max(np.abs(got - ref)) / scale
    H.close()


@pytest.mark.parametrize("nx,ny,nz,kind", [(20, 20, 20, "cube"), (21, 21, 21, "sphere"), (70, 12, 9, "cube"),
                                             (67, 67, 67, "sphere"), (106, 106, 106, "sphere")])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("nvec", [9, 17])
def test_stencil_compacted_columns_match_lean(gpu, nx, ny, nz, kind, dtype, nvec):
    """The compacted-column z-march (lanes dealt the slab's actual columns)
    gives the same bits as the fixed x-chunk one, for the recurrence step and
    the first step (no X), radius 4."""
    import torch
    from parsec_python_amd import _lib
    from parsec_python_amd.hamiltonian import Hamiltonian, leading_dim
    print(f"stencil cc case {nx}x{ny}x{nz} {kind} {dtype} nvec={nvec}", flush=True)
    g = _grid(nx, ny, nz, 0.45, kind)
    n = g.npts
    rng = np.random.default_rng(nx + 5 * nvec)
    tdt = torch.float64 if dtype == "f64" else torch.float32
    H = Hamiltonian(g, 4, None, torch.tensor(rng.standard_normal(n), device="cuda"), tdt)
    ld = leading_dim(n)
    Y = torch.zeros((nvec, ld), dtype=tdt, device="cuda")
    X = torch.zeros((nvec, ld), dtype=tdt, device="cuda")
    Y[:, :n] = torch.tensor(rng.standard_normal((nvec, n)), dtype=tdt)
    X[:, :n] = torch.tensor(rng.standard_normal((nvec, n)), dtype=tdt)
    outs = {}
    try:
        for variant in (0, 1):
            _lib.call("pc_stencil_set_variant", variant)
            o1 = torch.zeros_like(X)
            o2 = torch.zeros_like(X)
            H.step(Y, X, o1, nvec, 0.7, -0.3, -1.1)
            H.step(Y, None, o2, nvec, 0.7, -0.3, 0.0)
            outs[variant] = (o1[:, :n].clone(), o2[:, :n].clone())
    finally:
        _lib.call("pc_stencil_set_variant", 2)
    assert torch.equal(outs[0][0], outs[1][0])
    assert torch.equal(outs[0][1], outs[1][1])
    H.close()


@pytest.mark.parametrize("entry", ["pc_stencil_f64", "pc_stencil_direct_f64"])
def test_stencil_dot_partials(gpu, entry):
    """Per-workgroup partial sums of out*y (vector-major) add up to the oracle's dots."""
    import ctypes as C
    import torch
    from parsec_python_amd import _lib
    from parsec_python_amd.device import ptr, stream
    from parsec_python_amd.grid import DeviceGrid
    from parsec_python_amd.hamiltonian import leading_dim
    g = _grid(21, 21, 21, 0.45, "sphere")
    n, nvec, radius = g.npts, 9, 4
    print(f"dot partials {entry}", flush=True)
    rng = np.random.default_rng(3)
    V = rng.standard_normal(n)
    Y = rng.standard_normal((nvec, n))
    ld = leading_dim(n)
    dg = DeviceGrid(g, radius)
    Yd = torch.zeros((nvec, ld), dtype=torch.float64, device="cuda")
    Yd[:, :n] = torch.tensor(Y)
    Od = torch.zeros_like(Yd)
    Vd = torch.tensor(V, device="cuda")
    if entry == "pc_stencil_f64":
        cnt, per = C.c_int(), C.c_int()
        _lib.call("pc_stencil_num_partials", dg.handle, 8, nvec, C.byref(cnt), C.byref(per))
        nper = per.value
    else:
        nper = (g.ny * g.nz + 3) // 4
    part = torch.zeros(nvec * nper, dtype=torch.float64, device="cuda")
    _lib.call(entry, dg.handle, ptr(Yd), C.c_void_p(0), ptr(Od), ptr(Vd), nvec, ld, 0.5 / 0.45 ** 2, 1.0, 0.0, 0.0,
              ptr(part), stream())
    torch.cuda.synchronize()
    L = _oracle_lap(g, radius) * (0.5 / 0.45 ** 2)
    ref_out = (L @ Y.T).T + V * Y
    dots = part.view(nvec, nper).sum(dim=1).cpu().numpy()
    assert np.allclose(dots, np.sum(ref_out * Y, axis=1), rtol=1e-11)
    dg.close()


def test_projectors_vs_oracle(gpu):
    import torch
    from parsec_python_amd import ionic, molecules as M
    from parsec_python_amd.grid import make_grid
    from parsec_python_amd.hamiltonian import Hamiltonian, leading_dim
    atoms, dom, _ = M.prepare(M.h2o_sample(), h_ang=0.3, radius_ang=4.0)
    for kind in ("cube", "sphere"):
        g = make_grid(dom["nx"], dom["h"], dom["radius"], kind)
        proj = ionic.DeviceProjectors(g, atoms)
        tabs = opseudo.projector_tables(dom, atoms, g.mask)
        assert proj.natoms == len(tabs)
        for t, (lo, hi) in zip(tabs, zip(proj.tables[0][:-1], proj.tables[0][1:])):
            assert np.array_equal(t["idx"], proj.tables[1][lo:hi])
            assert np.allclose(t["u"], proj.tables[2][lo:hi], rtol=1e-13, atol=1e-16)
        n = g.npts
        nvec = 5
        rng = np.random.default_rng(7)
        Y = rng.standard_normal((nvec, n))
        H = Hamiltonian(g, 4, proj, torch.zeros(n, dtype=torch.float64, device="cuda"), torch.float64)
        ld = leading_dim(n)
        Yd = torch.zeros((nvec, ld), dtype=torch.float64, device="cuda")
        Yd[:, :n] = torch.tensor(Y)
        out = torch.zeros_like(Yd)
        H.apply(Yd, out, nvec)
        L = ofd.laplacian_masked(g.nx, g.ny, g.nz, 4, g.mask) * (0.5 / g.h ** 2)
        ref = (L @ Y.T).T + opseudo.apply_nl(tabs, Y.T).T
        got = out[:, :n].cpu().numpy()
        assert np.max(np.abs(got - ref)) <= 1e-11 * np.abs(ref).max()
        H.close()
        proj.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("n,k", [(1000, 23), (5003, 49), (20000, 70), (513, 1), (3001, 275), (1201, 300)])
def test_gram_and_rotate(gpu, dtype, n, k):
    """S = X^T X, G = X^T Y and Z = X Q against NumPy.  The cases cover one to
    nine 32-column groups of the all-columns FP32 rotation (k + 3 <= 288), its
    fallback above that, ragged point counts and partial 16-vector slabs."""
    import torch
    from parsec_python_amd.hamiltonian import BlockLinalg, leading_dim
    print(f"gram case n={n} k={k} {dtype}", flush=True)
    tdt = torch.float64 if dtype == "f64" else torch.float32
    rng = np.random.default_rng(n + k)
    X = rng.standard_normal((k, n))
    Y = rng.standard_normal((k + 3, n))
    ld = leading_dim(n)
    Xd = torch.zeros((k, ld), dtype=tdt, device="cuda")
    Yd = torch.zeros((k + 3, ld), dtype=tdt, device="cuda")
    Xd[:, :n] = torch.tensor(X, dtype=tdt)
    Yd[:, :n] = torch.tensor(Y, dtype=tdt)
    Xr = Xd[:, :n].double().cpu().numpy()
    Yr = Yd[:, :n].double().cpu().numpy()
    la = BlockLinalg(n, k + 3)
    S = torch.empty((k, k), dtype=torch.float64, device="cuda")
    la.gram(Xd, Xd, k, k, S.view(-1), sym=True)
    G = torch.empty((k + 3, k), dtype=torch.float64, device="cuda")
    la.gram(Xd, Yd, k, k + 3, G.view(-1), sym=False)
    Sref = Xr @ Xr.T
    Gref = Xr @ Yr.T  # [i, j] = X_i . Y_j ; stored column-major (j-major rows)
    tol = 1e-12 if dtype == "f64" else 3e-6
    assert np.max(np.abs(S.cpu().numpy() - Sref)) <= tol * np.abs(Sref).max() * (1 if dtype == "f64" else 10)
    assert np.max(np.abs(G.cpu().numpy().T - Gref)) <= tol * np.abs(Sref).max() * (1 if dtype == "f64" else 10)
    Q = rng.standard_normal((k + 3, k))  # column-major k x (k+3) -> row j = column j
    Qd = torch.tensor(Q, device="cuda")
    Zref = (Xr.T @ Q.T).T
    ztol = 1e-12 if dtype == "f64" else 2e-6
    # FP32: the FP32 MFMA rotation (default) and the optional split-bf16 one, both to the FP32 bar
    from parsec_python_amd import _lib
    for split in ((0, 1) if dtype == "f32" else (0,)):
        _lib.call("pc_rotate_set_split_bf16", split)
        Z = torch.zeros((k + 3, ld), dtype=tdt, device="cuda")
        la.rotate(Xd, Qd.view(-1), k, k + 3, Z)
        err = np.max(np.abs(Z[:, :n].double().cpu().numpy() - Zref))
        print(f"  rotate split_bf16={split}: max error {err / np.abs(Zref).max():.2e} relative", flush=True)
        assert err <= ztol * np.abs(Zref).max()
    _lib.call("pc_rotate_set_split_bf16", 0)
    la.close()


def test_dense_rr_generalized_eig(gpu):
    import torch
    from scipy.linalg import eigh
    from parsec_python_amd.hamiltonian import BlockLinalg
    rng = np.random.default_rng(5)
    k = 40
    B = rng.standard_normal((200, k))
    S = B.T @ B
    M = rng.standard_normal((k, k))
    G = M + M.T
    la = BlockLinalg(10, k)
    Sd = torch.tensor(S.T.copy(), device="cuda").view(-1)
    Gd = torch.tensor(G.T.copy(), device="cuda").view(-1)
    T, lam = la.dense_rr(Sd, Gd, k)
    ref = eigh(G, S, eigvals_only=True)
    assert np.allclose(lam, ref, rtol=1e-10, atol=1e-10)
    Tm = T.cpu().numpy().reshape(k, k).T  # column-major
    assert np.allclose(Tm.T @ S @ Tm, np.eye(k), atol=1e-9)
    assert np.allclose(Tm.T @ G @ Tm, np.diag(lam), atol=1e-8)
    la.close()


@pytest.mark.parametrize("method", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("k,kind", [(1, "rand"), (2, "rand"), (3, "rand"), (33, "rand"), (65, "rand"),
                                    (128, "rand"), (275, "degenerate"), (275, "band"), (300, "rand"),
                                    (320, "degenerate"), (400, "rand"), (520, "rand")])
def test_dense_rr_methods(gpu, method, k, kind):
    """Every eigensolver of the dense stage (syevd, syevj, tridiagonalization
    with 32 workgroups (methods 2 and 4, 128 <= k <= 320) or one (method 3),
    then this library's tridiagonal eigensolver (2, 3) or rocSOLVER's stedc
    (4)) returns the generalized eigenpairs of (G, S): T^T S T = I,
    T^T G T = diag(lam), lam = SciPy's.  'degenerate' has exact triple
    eigenvalues, as the symmetric clusters give; 'band' packs 3/4 of the
    spectrum into a narrow window (many close pairs), as the occupied Ritz
    values of a large cluster do."""
    import torch
    from scipy.linalg import eigh
    from parsec_python_amd import _lib
    from parsec_python_amd.hamiltonian import BlockLinalg
    if method == 1 and k > 300:
        pytest.skip("Jacobi at this size is slow and covered at k <= 300")
    print(f"dense_rr method {method} k {k} {kind}", flush=True)
    rng = np.random.default_rng(k + 7 * method)
    B = rng.standard_normal((max(2 * k, 4), k))
    S = B.T @ B / B.shape[0] + 0.5 * np.eye(k)
    if kind in ("degenerate", "band"):
        Qr, _ = np.linalg.qr(rng.standard_normal((k, k)))
        if kind == "degenerate":
            lam0 = np.repeat(rng.uniform(-1.0, 2.0, (k + 2) // 3), 3)[:k]
        else:
            lam0 = np.sort(np.concatenate([rng.uniform(-0.5, -0.3, 3 * k // 4), rng.uniform(0.0, 0.3, k - 3 * k // 4)]))
        A = (Qr * lam0) @ Qr.T
        L = np.linalg.cholesky(S)
        G = L @ A @ L.T                      # (G, S) has the spectrum lam0
    else:
        M = rng.standard_normal((k, k))
        G = M + M.T
    la = BlockLinalg(10, k)
    _lib.call("pc_dense_set_method", la.dense, method)
    ref = eigh(G, S, eigvals_only=True)
    scale = max(1.0, np.abs(ref).max())
    # R^-1 from the blocked Cholesky kernels (the default; rocSOLVER's potrf
    # + trtri above k = 320), and (method 2) also as trtri + trmm and as
    # triangular solves; Q Z (methods 2-4) by this library's kernels (the
    # default, k <= 320) and by rocSOLVER's ormtr
    combos = [(2, 1), (0, 1), (1, 1)] if method == 2 else [(2, 1)]
    if method in (2, 3, 4):
        combos.append((2, 0))
    for tri_mode, orm in combos:
        _lib.call("pc_dense_set_triangular", la.dense, tri_mode)
        _lib.call("pc_dense_set_backtransform", la.dense, orm)
        Sd = torch.tensor(S.T.copy(), device="cuda").view(-1)
        Gd = torch.tensor(G.T.copy(), device="cuda").view(-1)
        T, lam = la.dense_rr(Sd, Gd, k)
        assert np.allclose(lam, ref, rtol=0, atol=1e-11 * scale * k)
        Tm = T.cpu().numpy()[:k * k].reshape(k, k).T  # column-major
        assert np.abs(Tm.T @ S @ Tm - np.eye(k)).max() < 1e-10 * k
        assert np.abs(Tm.T @ G @ Tm - np.diag(lam)).max() < 1e-10 * scale * k
    aborts = C.c_int(-1)
    _lib.call("pc_dense_trd_control", la.dense, 0, 0, C.byref(aborts))
    assert aborts.value == 0          # the multi-workgroup waits all completed
    la.close()


@pytest.mark.parametrize("k", [128, 275])
def test_dense_rr_mwg_fallback(gpu, k):
    """With a one-poll limit on the multi-workgroup tridiagonalization's waits
    (forced from the host; every workgroup still exits), pc_dense_rr repeats
    the reduction with one workgroup and the result is unchanged."""
    import torch
    from scipy.linalg import eigh
    from parsec_python_amd import _lib
    from parsec_python_amd.hamiltonian import BlockLinalg
    print(f"dense_rr fallback k {k}", flush=True)
    rng = np.random.default_rng(31 + k)
    B = rng.standard_normal((2 * k, k))
    S = B.T @ B / B.shape[0] + 0.5 * np.eye(k)
    M = rng.standard_normal((k, k))
    G = M + M.T
    ref = eigh(G, S, eigvals_only=True)
    la = BlockLinalg(10, k)
    _lib.call("pc_dense_trd_control", la.dense, 1, 0, None)
    aborts = C.c_int(-1)
    for rep in range(2):
        Sd = torch.tensor(S.T.copy(), device="cuda").view(-1)
        Gd = torch.tensor(G.T.copy(), device="cuda").view(-1)
        T, lam = la.dense_rr(Sd, Gd, k)
        assert np.allclose(lam, ref, rtol=0, atol=1e-11 * max(1.0, np.abs(ref).max()) * k)
        Tm = T.cpu().numpy()[:k * k].reshape(k, k).T
        assert np.abs(Tm.T @ S @ Tm - np.eye(k)).max() < 1e-10 * k
    _lib.call("pc_dense_trd_control", la.dense, 0, 0, C.byref(aborts))
    print(f"  reductions repeated with one workgroup: {aborts.value}", flush=True)
    assert aborts.value in (0, 1)     # at most once: the context then keeps one workgroup
    _lib.call("pc_dense_trd_control", la.dense, 1 << 21, 1, None)
    la.close()


def test_dense_rr_large_cluster_falls_back_to_stedc(gpu):
    """A cluster of 40 eigenvalues within 1e-7 of each other is more than the
    tridiagonal eigensolver's clusters take (32): the step is repeated with
    rocSOLVER's stedc once, the context keeps stedc, and the pairs are right."""
    import torch
    from scipy.linalg import eigh
    from parsec_python_amd import _lib
    from parsec_python_amd.hamiltonian import BlockLinalg
    k = 128
    rng = np.random.default_rng(77)
    Qr, _ = np.linalg.qr(rng.standard_normal((k, k)))
    lam0 = np.sort(np.concatenate([0.2 + 1e-7 * rng.standard_normal(40), rng.uniform(-1.0, 1.0, k - 40)]))
    G = (Qr * lam0) @ Qr.T
    G = 0.5 * (G + G.T)
    S = np.eye(k)
    la = BlockLinalg(10, k)
    fb = C.c_int(-1)
    for rep in range(2):
        T, lam = la.dense_rr(torch.tensor(S.T.copy(), device="cuda").view(-1),
                             torch.tensor(G.T.copy(), device="cuda").view(-1), k)
        ref = eigh(G, S, eigvals_only=True)
        assert np.allclose(lam, ref, rtol=0, atol=1e-12 * k)
        Tm = T.cpu().numpy()[:k * k].reshape(k, k).T
        assert np.abs(Tm.T @ Tm - np.eye(k)).max() < 1e-10 * k
        assert np.abs(Tm.T @ G @ Tm - np.diag(lam)).max() < 1e-10 * k
    _lib.call("pc_dense_tri_fallbacks", la.dense, C.byref(fb))
    print(f"  steps repeated with stedc: {fb.value}", flush=True)
    assert fb.value == 1
    la.close()


def test_dense_rr_timing_tridiagonal_eigensolvers(gpu):
    """Time of one Rayleigh-Ritz dense stage at k = 275 with this library's
    tridiagonal eigensolver (method 2) and with rocSOLVER's stedc (method 4),
    with R^-1 from rocSOLVER's potrf + trtri, from triangular solves and from
    the blocked Cholesky kernels, and Q Z by this library's kernels or by
    rocSOLVER's ormtr; all give the same pairs.  Printed for the record; the
    bars are that the library's eigensolver is not slower than stedc, the
    blocked Cholesky kernels not slower than potrf + trtri and the library's
    Q Z not slower than ormtr."""
    import time
    import torch
    from parsec_python_amd import _lib
    from parsec_python_amd.hamiltonian import BlockLinalg
    k = 275
    rng = np.random.default_rng(5)
    B = rng.standard_normal((2 * k, k))
    S = B.T @ B / B.shape[0] + 0.5 * np.eye(k)
    M = rng.standard_normal((k, k))
    G = M + M.T
    Sd0 = torch.tensor(S.T.copy(), device="cuda").view(-1)
    Gd0 = torch.tensor(G.T.copy(), device="cuda").view(-1)
    out = {}
    for method, tri_mode, orm in ((2, 0, 1), (4, 0, 1), (2, 1, 1), (2, 2, 1), (2, 2, 0)):
        la = BlockLinalg(10, k)
        _lib.call("pc_dense_set_method", la.dense, method)
        _lib.call("pc_dense_set_triangular", la.dense, tri_mode)
        _lib.call("pc_dense_set_backtransform", la.dense, orm)
        ts = []
        for rep in range(6):
            Sd, Gd = Sd0.clone(), Gd0.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            T, lam = la.dense_rr(Sd, Gd, k)
            ts.append(time.perf_counter() - t0)
        out[(method, tri_mode) if orm else "ormtr"] = (float(np.median(ts[1:])) * 1e3, lam)
        la.close()
    print(f"  dense stage k {k}: tridiagonal eigensolver {out[(2, 0)][0]:.2f} ms, stedc {out[(4, 0)][0]:.2f} ms "
          f"(potrf + trtri); with triangular solves {out[(2, 1)][0]:.2f} ms; with the blocked Cholesky kernels "
          f"{out[(2, 2)][0]:.2f} ms, and with rocSOLVER's ormtr for Q Z {out['ormtr'][0]:.2f} ms", flush=True)
    for key in ((2, 0), (2, 1), (2, 2), "ormtr"):
        assert np.allclose(out[key][1], out[(4, 0)][1], rtol=0, atol=1e-12 * k * np.abs(out[(4, 0)][1]).max())
    assert out[(2, 0)][0] <= 1.05 * out[(4, 0)][0]
    assert out[(2, 2)][0] <= 1.05 * out[(2, 0)][0]
    assert out[(2, 2)][0] <= 1.05 * out["ormtr"][0]


@pytest.mark.parametrize("k", [34, 275])
def test_dense_rr_with_cholesky_on_the_second_stream(gpu, k):
    """The Cholesky factor queued on the dense stage's second stream
    (pc_dense_chol_async) while other work runs on the main stream gives
    pc_dense_rr the same bits as computing it in place; a factor queued for a
    different S is waited for and not used."""
    import torch
    from parsec_python_amd.hamiltonian import BlockLinalg
    rng = np.random.default_rng(11 + k)
    B = rng.standard_normal((2 * k, k))
    S = B.T @ B / B.shape[0] + 0.5 * np.eye(k)
    M = rng.standard_normal((k, k))
    G = M + M.T
    S0 = torch.tensor(S.T.copy(), device="cuda").view(-1)
    G0 = torch.tensor(G.T.copy(), device="cuda").view(-1)
    la = BlockLinalg(10, k)
    T1, lam1 = la.dense_rr(S0.clone(), G0.clone(), k)
    T1 = T1.clone()
    Sd, Gd = S0.clone(), G0.clone()
    la.dense_chol_async(Sd, k)
    busy = torch.randn(4096, 4096, device="cuda")   # main-stream work in between
    busy = busy @ busy
    T2, lam2 = la.dense_rr(Sd, Gd, k)
    assert torch.equal(T1, T2) and np.array_equal(lam1, lam2)
    other = S0 * 2.0                                 # kept alive: a different S at another address
    la.dense_chol_async(other, k)                    # not the S below: waited for, not used
    T3, lam3 = la.dense_rr(S0.clone(), G0.clone(), k)
    assert torch.equal(T1, T3) and np.array_equal(lam1, lam3)
    torch.cuda.synchronize()
    la.close()


@pytest.mark.parametrize("k", [34, 275])
def test_dense_rr_after_nan_filled_device_memory(gpu, k):
    """Device memory released by earlier work and handed to a new dense
    context holds whatever was there (here NaN): the Rayleigh-Ritz stage must
    read none of it before writing (the inverse iteration once read the last
    row's superdiagonals of its LU workspace unwritten)."""
    import torch
    from scipy.linalg import eigh
    from parsec_python_amd.hamiltonian import BlockLinalg
    junk = torch.full((96 * 1024 * 1024,), float("nan"), dtype=torch.float64, device="cuda")
    del junk
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    rng = np.random.default_rng(3 + k)
    B = rng.standard_normal((2 * k, k))
    S = B.T @ B / B.shape[0] + 0.5 * np.eye(k)
    M = rng.standard_normal((k, k))
    G = M + M.T
    la = BlockLinalg(10, k)
    T, lam = la.dense_rr(torch.tensor(S.T.copy(), device="cuda").view(-1),
                         torch.tensor(G.T.copy(), device="cuda").view(-1), k)
    Tm = T.cpu().numpy()[:k * k].reshape(k, k).T
    assert np.all(np.isfinite(lam)) and np.all(np.isfinite(Tm))
    assert np.allclose(lam, eigh(G, S, eigvals_only=True), rtol=0, atol=1e-11 * k * max(1.0, np.abs(lam).max()))
    assert np.abs(Tm.T @ S @ Tm - np.eye(k)).max() < 1e-10 * k
    la.close()


@pytest.mark.parametrize("tri_mode", [0, 2])
@pytest.mark.parametrize("k", [150, 275])
def test_dense_rr_reports_indefinite_overlap(gpu, tri_mode, k):
    """An overlap matrix with a negative pivot at column 101 that no shift of
    at most 1e-7 of its largest diagonal entry repairs: both Cholesky forms
    (rocSOLVER's potrf and the blocked kernels, whose failed attempt is
    repeated with potrf) report the same column, and pc_dense_rr fails with
    that message after its four shifted retries."""
    import torch
    from parsec_python_amd import _lib
    from parsec_python_amd._lib import ParsecHipError
    from parsec_python_amd.hamiltonian import BlockLinalg
    S = np.eye(k)
    S[100, 100] = -1.0
    G = np.diag(np.arange(k, dtype=float))
    la = BlockLinalg(10, k)
    _lib.call("pc_dense_set_triangular", la.dense, tri_mode)
    with pytest.raises(ParsecHipError, match="potrf info 101"):
        la.dense_rr(torch.tensor(S.T.copy(), device="cuda").view(-1), torch.tensor(G.T.copy(), device="cuda").view(-1), k)
    la.close()


def test_density_xc_post_vs_oracle(gpu):
    import torch
    from parsec_python_amd import _lib
    from parsec_python_amd.device import ptr, stream
    rng = np.random.default_rng(11)
    n, nev, ld = 4099, 6, 4112
    W = rng.standard_normal((nev + 2, ld)) * 0.01
    occ2 = rng.uniform(0, 2, nev)
    Wd = torch.tensor(W, device="cuda")
    o2 = torch.tensor(occ2, device="cuda")
    rho_cell = torch.empty(n, dtype=torch.float64, device="cuda")
    _lib.call("pc_density_f64", ptr(Wd), ld, n, nev, ptr(o2), ptr(rho_cell), stream())
    ref = (W[:nev, :n] ** 2 * occ2[:, None]).sum(0)
    assert np.allclose(rho_cell.cpu().numpy(), ref, rtol=1e-13, atol=1e-18)
    # include zero / tiny / large densities (both PZ81 branches)
    rc = np.abs(rng.standard_normal(n)) * 0.05
    rc[:10] = 0.0
    rc[10:20] = 1e-12
    h3 = 0.3 ** 3
    ppot, hpot, hpot0, pot = (rng.standard_normal(n) for _ in range(4))
    dv = [torch.tensor(v, device="cuda") for v in (rc, ppot, hpot, hpot0, pot)]
    rho = torch.empty(n, dtype=torch.float64, device="cuda")
    vxc = torch.empty_like(rho)
    pn = torch.empty_like(rho)
    nb = C.c_int()
    _lib.call("pc_num_blocks", n, C.byref(nb))
    part = torch.empty(5 * nb.value, dtype=torch.float64, device="cuda")
    sums = torch.empty(5, dtype=torch.float64, device="cuda")
    _lib.call("pc_scf_post", n, ptr(dv[0]), h3, ptr(dv[1]), ptr(dv[2]), ptr(dv[3]), ptr(dv[4]), ptr(rho), ptr(vxc),
              ptr(pn), ptr(part), ptr(sums), stream())
    rr = rc / h3
    v_ref, e_ref = oxc.exc_nspn(1.0, rr)
    assert np.allclose(vxc.cpu().numpy(), v_ref, rtol=1e-13, atol=1e-15)
    pn_ref = ppot + 0.5 * v_ref + hpot + hpot0
    assert np.allclose(pn.cpu().numpy(), pn_ref, rtol=1e-14, atol=1e-14)
    s = sums.cpu().numpy()
    assert s[0] == pytest.approx(e_ref, rel=1e-12)
    assert s[1] == pytest.approx(np.sum((pn_ref - pot) ** 2), rel=1e-12)
    assert s[2] == pytest.approx(np.sum(pn_ref ** 2), rel=1e-12)
    assert s[3] == pytest.approx(np.sum(rr * (hpot + hpot0)), rel=1e-11)
    assert s[4] == pytest.approx(np.sum(rr * v_ref), rel=1e-12)


@pytest.mark.parametrize("kind", ["cube", "sphere"])
@pytest.mark.parametrize("alias", [True, False])
def test_pcg_vs_oracle(gpu, kind, alias):
    import torch
    from parsec_python_amd.grid import DeviceGrid, make_grid
    from parsec_python_amd.hartree import HartreeSolver
    h = 0.4
    g = make_grid(24, h, 0.5 * h * 23, kind)
    n = g.npts
    A = (1.0 / h ** 2) * ofd.laplacian_masked(g.nx, g.ny, g.nz, 4, g.mask)
    rng = np.random.default_rng(3)
    rhs = rng.standard_normal(n)
    x0 = rng.standard_normal(n) * 0.1
    xr, its_r = opcg(A, rhs, x0, 200, 1e-5, reference_aliasing=alias)
    dg = DeviceGrid(g, 4)
    hs = HartreeSolver(dg, h, reference_aliasing=alias, check_every=7)
    x = torch.tensor(x0, device="cuda")
    its = hs.solve(torch.tensor(rhs, device="cuda"), x, 200, 1e-5)
    assert its == its_r
    assert np.max(np.abs(x.cpu().numpy() - xr)) <= 1e-9 * np.abs(xr).max()
    # warm start (the stop rule is relative to the initial residual, pcg.py:50)
    x1 = xr + 1e-3 * rng.standard_normal(n)
    xr2, its_r2 = opcg(A, rhs, x1, 200, 1e-5, reference_aliasing=alias)
    x2 = torch.tensor(x1, device="cuda")
    its2 = hs.solve(torch.tensor(rhs, device="cuda"), x2, 200, 1e-5)
    assert its2 == its_r2
    assert np.max(np.abs(x2.cpu().numpy() - xr2)) <= 1e-9 * np.abs(xr2).max()
    dg.close()


@pytest.mark.parametrize("nx,ny,nz,kind,radius", [(70, 70, 70, "sphere", 4), (130, 9, 7, "cube", 6),
                                                  (40, 13, 22, "cube", 2), (66, 66, 66, "sphere", 1)])
def test_pcg_shapes_vs_oracle(gpu, nx, ny, nz, kind, radius):
    """Iteration counts and solutions on ragged and wide grids (x widths of 1, 2 and 4 wavefronts, slabs and
    z-chunks that do not divide the grid)."""
    import torch
    from parsec_python_amd.grid import DeviceGrid
    from parsec_python_amd.hartree import HartreeSolver
    print(f"pcg shape {nx}x{ny}x{nz} {kind} r={radius}", flush=True)
    h = 0.35
    g = _grid(nx, ny, nz, h, kind)
    n = g.npts
    A = (1.0 / h ** 2) * ofd.laplacian_masked(g.nx, g.ny, g.nz, radius, g.mask)
    rng = np.random.default_rng(nx + ny + nz)
    rhs = rng.standard_normal(n)
    x0 = np.zeros(n)
    xr, its_r = opcg(A, rhs, x0, 60, 1e-5, reference_aliasing=True)
    dg = DeviceGrid(g, radius)
    hs = HartreeSolver(dg, h, reference_aliasing=True, check_every=16)
    x = torch.tensor(x0, device="cuda")
    its = hs.solve(torch.tensor(rhs, device="cuda"), x, 60, 1e-5)
    assert its == its_r
    assert np.max(np.abs(x.cpu().numpy() - xr)) <= 1e-9 * np.abs(xr).max()
    dg.close()


@pytest.mark.parametrize("nx,ny,nz,kind,radius", [(24, 24, 24, "sphere", 4), (40, 13, 22, "cube", 2),
                                                  (70, 70, 70, "sphere", 4), (33, 33, 33, "cube", 6)])
def test_pcg_multigrid(gpu, nx, ny, nz, kind, radius):
    """Multigrid-preconditioned CG: the solution meets the same stop rule (true residual <= tol x the initial
    residual, checked against the oracle's sparse operator), matches a direct solve, and needs far fewer
    iterations than plain CG."""
    import torch
    from parsec_python_amd.grid import DeviceGrid
    from parsec_python_amd.hartree import HartreeSolver
    print(f"pcg mg {nx}x{ny}x{nz} {kind} r={radius}", flush=True)
    h = 0.35
    g = _grid(nx, ny, nz, h, kind)
    n = g.npts
    A = ((1.0 / h ** 2) * ofd.laplacian_masked(g.nx, g.ny, g.nz, radius, g.mask)).tocsr()
    rng = np.random.default_rng(nx * 7 + radius)
    # smooth Gaussian charges (the shape of a Hartree source) plus grid-scale noise
    pts = np.argwhere(g.mask.reshape(nz, ny, nx)).astype(np.float64)          # (k, j, i) of the domain points
    dims = np.array([nz, ny, nx], dtype=np.float64)
    rhs = rng.standard_normal(n)
    for _ in range(3):
        ctr = dims / 2 + rng.uniform(-0.15, 0.15, 3) * dims
        rhs += 40.0 * np.exp(-np.sum((pts - ctr) ** 2, axis=1) / (2 * (0.08 * dims.min()) ** 2))
    xd, _ = opcg(A, rhs, np.zeros(n), 5000, 1e-13, reference_aliasing=False)   # reference solution
    dg = DeviceGrid(g, radius)
    hs = HartreeSolver(dg, h, check_every=4, prec="mg")
    hp = HartreeSolver(dg, h, reference_aliasing=False, check_every=16)
    for x0 in (np.zeros(n), xd + 1e-3 * rng.standard_normal(n)):
        x = torch.tensor(x0, device="cuda")
        its = hs.solve(torch.tensor(rhs, device="cuda"), x, 500, 1e-5)
        xg = x.cpu().numpy()
        r0 = np.linalg.norm(rhs - A @ x0)
        assert np.linalg.norm(rhs - A @ xg) <= 1.05e-5 * r0
        x2 = torch.tensor(x0, device="cuda")
        its_plain = hp.solve(torch.tensor(rhs, device="cuda"), x2, 2000, 1e-5)
        print(f"   its mg {its} plain {its_plain}", flush=True)
        assert its < its_plain
        if n > 100000:
            assert 3 * its < its_plain
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    hs.solve(torch.tensor(rhs, device="cuda"), x, 500, 1e-10)
    assert np.max(np.abs(x.cpu().numpy() - xd)) <= 1e-7 * np.abs(xd).max()
    hs.close()
    dg.close()


def test_pcg_multigrid_check_schedule_gives_the_same_bits(gpu):
    """A warm solver (first state readback after about the previous solve's
    iteration count, then every 2) and a fresh one (every check_every) give
    the same iterations and the same solution bits: kernels queued past
    convergence return at once."""
    import torch
    from parsec_python_amd.grid import DeviceGrid
    from parsec_python_amd.hartree import HartreeSolver
    h = 0.35
    g = _grid(40, 40, 40, h, "sphere")
    n = g.npts
    rng = np.random.default_rng(41)
    rhs1, rhs2 = (torch.tensor(rng.standard_normal(n), device="cuda") for _ in range(2))
    dg = DeviceGrid(g, 4)
    warm = HartreeSolver(dg, h, prec="mg")
    fresh = HartreeSolver(dg, h, prec="mg")
    warm.solve(rhs1, torch.zeros(n, dtype=torch.float64, device="cuda"), 500, 1e-5)
    print(f"   first solve {warm.last_its} iterations", flush=True)
    for tol in (1e-5, 1e-7):
        xa = torch.zeros(n, dtype=torch.float64, device="cuda")
        xb = torch.zeros(n, dtype=torch.float64, device="cuda")
        ia = warm.solve(rhs2, xa, 500, tol)
        ib = fresh.solve(rhs2, xb, 500, tol)
        assert ia == ib
        assert torch.equal(xa, xb)
    warm.close()
    fresh.close()
    dg.close()


@pytest.mark.parametrize("kind", ["cube", "sphere"])
def test_sad_eval_vs_oracle(gpu, kind):
    from parsec_python_amd import ionic, molecules as M
    from parsec_python_amd.grid import DeviceGrid, make_grid
    atoms, dom, _ = M.prepare(M.h2o_sample(), h_ang=0.2, radius_ang=5.0)
    g = make_grid(dom["nx"], dom["h"], dom["radius"], kind)
    dg = DeviceGrid(g, 4)
    rho0, hpot0, pot, info = ionic.pseudo_diag(g, dg, atoms)
    r0, h0, p0, i0 = opseudo.pseudo_diag(dom, atoms, g.mask)
    for name, got, ref in (("rho0", rho0, r0), ("hpot0", hpot0, h0), ("pot", pot, p0)):
        d = np.abs(got.cpu().numpy() - ref)
        print(f"sad {kind} {name}: max|d|={d.max():.3e} at {d.argmax()} got={got.cpu().numpy()[d.argmax()]:.6e}"
              f" ref={ref[d.argmax()]:.6e} max|ref|={np.abs(ref).max():.3e}", flush=True)
    print("info", info, i0, flush=True)
    assert np.allclose(rho0.cpu().numpy(), r0, rtol=1e-12, atol=1e-16)
    assert np.allclose(hpot0.cpu().numpy(), h0, rtol=1e-12, atol=1e-14)
    assert np.allclose(pot.cpu().numpy(), p0, rtol=1e-12, atol=1e-14)
    assert info["electron_count_initial"] == pytest.approx(i0["electron_count_initial"], rel=1e-13)
    if kind == "cube":
        gold = np.load("tests/golden/h2o_sample_rho_grid.npz")["init"]
        assert np.max(np.abs(rho0.cpu().numpy() - gold)) < 1e-15
    dg.close()


def test_rand_matches_mirror(gpu):
    import torch
    from parsec_python_amd.eigensolvers import rand_block
    from tests.rng_mirror import block
    n, k = 1001, 3
    out = torch.empty((k, 1008), dtype=torch.float64, device="cuda")
    rand_block(out, n, k, 12345, normal=False)
    assert np.array_equal(out[:, :n].cpu().numpy(), block(12345, n, k, False).T)
    rand_block(out, n, k, 777, normal=True)
    assert np.allclose(out[:, :n].cpu().numpy(), block(777, n, k, True).T, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("n", [1, 3000, 70001])
def test_lanczos_step_vs_numpy(gpu, n):
    """pc_lanczos_start / pc_lanczos_step: the three-term recurrence with alpha and
    beta kept on the device equals NumPy's on the same start vector, given
    w = M v1 from the host each step (M symmetric)."""
    import ctypes as C
    import torch
    from parsec_python_amd import _lib
    from parsec_python_amd.device import ptr, stream
    print(f"lanczos step n={n}", flush=True)
    rng = np.random.default_rng(n)
    m = 6 if n > 1 else 1
    if n <= 3000:
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        lam = rng.uniform(-3.0, 5.0, n)
        matvec = lambda x: Q @ (lam * (Q.T @ x))
    else:
        d = rng.uniform(-3.0, 5.0, n)          # a diagonal operator keeps the big case cheap
        matvec = lambda x: d * x
    v = rng.standard_normal(n)
    nb = C.c_int()
    _lib.call("pc_lanczos_nblocks", n, C.byref(nb))
    part = torch.empty(2 * nb.value, dtype=torch.float64, device="cuda")
    coef = torch.zeros(2 * m + 2, dtype=torch.float64, device="cuda")
    v0 = torch.zeros(n, dtype=torch.float64, device="cuda")
    v1 = torch.tensor(v, device="cuda")
    w = torch.empty(n, dtype=torch.float64, device="cuda")
    _lib.call("pc_lanczos_start", n, ptr(v1), ptr(part), stream())
    # NumPy reference with the same arithmetic: w = M v1; alpha = v1.w; w -= alpha v1 + beta v0
    r1 = v / np.linalg.norm(v)
    r0 = np.zeros(n)
    beta = 0.0
    ref = []
    for k in range(m):
        w.copy_(torch.tensor(matvec(v1.cpu().numpy()), device="cuda"))
        _lib.call("pc_lanczos_step", n, ptr(v0), ptr(v1), ptr(w), ptr(coef), k, ptr(part), stream())
        v0, v1, w = v1, w, v0
        rw = matvec(r1)
        alpha = r1 @ rw
        rw = rw - alpha * r1 - beta * r0
        beta = np.linalg.norm(rw)
        ref += [alpha, beta]
        r0, r1 = r1, rw / beta if beta > 0 else rw
        if n == 1:
            break
    got = coef.cpu().numpy()[:len(ref)]
    if n == 1:
        assert abs(got[0] - ref[0]) <= 1e-12 * max(1.0, abs(ref[0]))
        return
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-12)


def test_lanczos_bound_is_upper_bound(gpu):
    """Eigensolver.lanczos_bound (device-side Lanczos) bounds the spectrum of H
    from above, within 20 % of the largest eigenvalue (oracle sparse H)."""
    import scipy.sparse as sps
    import scipy.sparse.linalg as spla
    import torch
    from parsec_python_amd.eigensolvers import Eigensolver
    from parsec_python_amd.hamiltonian import Hamiltonian
    h = 0.4
    g = _grid(22, 22, 22, h, "sphere")
    n = g.npts
    V = np.random.default_rng(3).uniform(-1.0, 0.5, n)
    H = Hamiltonian(g, 4, None, torch.tensor(V, device="cuda"), torch.float64)
    eig = Eigensolver(H, 8, seed=5)
    bound = eig.lanczos_bound(6, 6)
    A = (0.5 / h ** 2) * _oracle_lap(g, 4) + sps.diags(V)
    lmax = spla.eigsh(A.tocsc(), k=1, which="LA", return_eigenvectors=False)[0]
    assert lmax <= bound <= 1.2 * lmax, (bound, lmax)
    eig.close()
    H.close()
