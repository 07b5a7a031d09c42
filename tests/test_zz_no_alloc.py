"""The no-allocation CG entry point (pc_pcg_solve_noalloc) against the
workspace form (pc_pcg_solve): the same iterations and the same solution bit
for bit, on two calls in a row, the second on a smaller grid (so stale scratch
contents from the first call are present), plus the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve_both(g, radius, h, seed, alias):
    import torch
    from parsec_python_amd import _lib
    from parsec_python_amd.device import ptr, stream
    from parsec_python_amd.grid import DeviceGrid
    dg = DeviceGrid(g, radius)
    n = dg.npts
    rng = np.random.default_rng(seed)
    rhs = torch.tensor(rng.standard_normal(n), device="cuda")
    x0 = rng.standard_normal(n) * 0.1
    nb = C.c_size_t()
    _lib.call("pc_pcg_workspace", dg.handle, C.byref(nb))
    work = torch.empty(nb.value // 8 + 1, dtype=torch.float64, device="cuda")
    out = {}
    for name in ("pc_pcg_solve", "pc_pcg_solve_noalloc"):
        x = torch.tensor(x0, device="cuda")
        its = C.c_int()
        res = (C.c_double * 2)()
        if name == "pc_pcg_solve":
            _lib.call(name, dg.handle, 1.0 / h ** 2, ptr(rhs), ptr(x), 300, 1e-6, alias, 5, ptr(work), nb.value,
                      C.byref(its), res, stream())
        else:
            _lib.call(name, dg.handle, 1.0 / h ** 2, ptr(rhs), ptr(x), 300, 1e-6, alias, 5, C.byref(its), res,
                      stream())
        torch.cuda.synchronize()
        out[name] = (its.value, x.cpu().numpy(), (res[0], res[1]))
    dg.close()
    return out


@pytest.mark.parametrize("alias", [0, 1])
def test_noalloc_matches_workspace_form(gpu, alias):
    from parsec_python_amd.grid import make_grid
    h = 0.35
    big = make_grid(60, h, 0.5 * h * 59, "sphere")
    small = make_grid(26, h, 0.5 * h * 25, "cube")
    for g, radius, seed in ((big, 4, 11), (small, 6, 12)):
        print(f"noalloc CG: {g.kind} {g.nx}^3 ({g.npts} points) r={radius} alias={alias}", flush=True)
        out = _solve_both(g, radius, h, seed, alias)
        its_a, xa, ra = out["pc_pcg_solve"]
        its_b, xb, rb = out["pc_pcg_solve_noalloc"]
        assert its_a == its_b and its_a > 5
        assert np.array_equal(xa, xb)
        assert ra == rb


def test_noalloc_rejects_grids_above_the_limit(gpu):
    import torch
    from parsec_python_amd import _lib
    from parsec_python_amd.device import ptr, stream
    from parsec_python_amd.grid import DeviceGrid, Grid
    lim = C.c_longlong()
    _lib.call("pc_pcg_noalloc_max_points", C.byref(lim))
    assert lim.value == 1 << 21
    nx, ny, nz = 130, 130, 125           # 2,112,500 points > 2^21
    g = Grid(nx=nx, ny=ny, nz=nz, h=0.3, rad=0.15 * (nx - 1), mask=np.ones(nx * ny * nz, bool), kind="cube")
    dg = DeviceGrid(g, 4)
    rhs = torch.zeros(dg.npts, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(rhs)
    its = C.c_int()
    st = _lib.lib().pc_pcg_solve_noalloc(dg.handle, 1.0, ptr(rhs), ptr(x), 10, 1e-5, 0, 4, C.byref(its), None,
                                         stream())
    assert st == 4
    assert b"exceeds the maximum" in _lib.lib().pc_last_error()
    dg.close()
