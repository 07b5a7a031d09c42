"""The drop-in backend driven exactly as the reference's SCF driver drives its
GPU bundle (src/rsdft_solver.py:255-352: B = 0.5*A + vnl + sp.diags(pot),
first_filt / chefsi1 / occupations on backend arrays, rho = W[:, :nev]**2 @ 2occ,
pcg(A, hrhs, Hpot, 200, 1e-5), XC and the mixer on the host), compared with
the FP64 oracle's run of the same loop.  Both draw their random start vectors
from the same counter-based stream, so eigenvalues and CG counts agree."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fd as ofd
from oracle import scf as oscf
from oracle import xc as oxc
from oracle.mixer import Msecant1
from tests.rng_mirror import MirrorRNG

pytestmark = pytest.mark.gpu


def test_backend_reference_driver_matches_oracle(gpu):
    from parsec_python_amd import molecules as M
    from parsec_python_amd.backend import make_backend
    atoms_t, dom, nev = M.prepare(M.h2o(), h_ang=0.3, radius_ang=4.0, nev=8)
    atoms = [{"typ": s, "coord": np.asarray(x)} for s, x in atoms_t]
    seed, maxits, poldeg, temp = 5, 6, 10, 500.0
    print(f"backend driver nx={dom['nx']}", flush=True)
    be = make_backend("fp64", seed=seed)
    h = dom["h"]
    rho0, hpot0, ppot, info = be.pseudo_diag(dom, atoms, None, None, return_info=True)
    vnl = be.pseudo_nl(dom, atoms, None, None)
    A = (1.0 / h ** 2) * ofd.laplacian(dom["nx"], dom["ny"], dom["nz"], 4)
    n = A.shape[0]
    st = oscf.build_setup(atoms_t, dom, "cube", 8)
    assert np.allclose(rho0, st.rho0, rtol=1e-10, atol=1e-14)
    assert np.allclose(hpot0, st.hpot0, rtol=1e-10, atol=1e-12)
    assert np.allclose(ppot, st.ppot, rtol=1e-10, atol=1e-12)
    ref = oscf.run_scf(st, nev, tol=1e-12, maxits=maxits, poldeg=poldeg, rng=MirrorRNG(seed))

    hpot = np.zeros(n)
    vxc, _ = oxc.exc_nspn(h, rho0 / h ** 3)
    pot = ppot + hpot0 + 0.5 * vxc
    mixer = Msecant1()
    half = 0.5 * A + vnl
    W = lam = None
    for it in range(1, maxits + 1):
        B = half + sp.diags(pot, 0, shape=(n, n))
        if it == 1:
            W, lam = be.first_filt(nev + 15, B, poldeg)
        else:
            W, lam = be.chefsi1(W, lam, poldeg, nev, B)
        _, occ = be.occupations(lam[:nev], temp, st.nelec, 1e-6)
        lam_host = np.asarray(be.to_numpy_array(lam)).reshape(-1)
        rho = (W[:, :nev] ** 2) @ (2 * occ)
        rho = np.asarray(be.to_numpy_array(rho)).reshape(-1)
        hrhs = (4 * np.pi / h ** 3) * (rho - rho0)
        rho = rho / h ** 3
        hpot, its = be.pcg(A, hrhs, hpot, 200, 1e-5)
        hpot = np.asarray(be.to_numpy_array(hpot)).reshape(-1)
        vxc, _ = oxc.exc_nspn(h, rho)
        pot_new = ppot + 0.5 * vxc + hpot + hpot0
        hr = ref.history[it - 1]
        assert np.allclose(lam_host[:nev], hr["lam"], atol=1e-9), it
        assert its == hr["cg_its"], it
        pot, _ = mixer(pot, pot_new - pot)
    assert np.allclose(pot, ref.pot, rtol=1e-8, atol=1e-10)


def test_backend_rejects_foreign_operators(gpu):
    from parsec_python_amd import molecules as M
    from parsec_python_amd.backend import make_backend
    atoms, dom, nev = M.prepare(M.h2o(), h_ang=0.3, radius_ang=4.0, nev=8)
    be = make_backend("fp64")
    be.pseudo_diag(dom, atoms, None, None)
    vnl = be.pseudo_nl(dom, atoms, None, None)
    h = dom["h"]
    A = (1.0 / h ** 2) * ofd.laplacian(dom["nx"], dom["ny"], dom["nz"], 4)
    with pytest.raises(ValueError):
        _ = 0.7 * A + vnl                      # not 0.5*A
    with pytest.raises(ValueError):
        be.first_filt(23, sp.identity(A.shape[0], format="csr"), 10)
    B = 0.5 * A + vnl + sp.diags(np.zeros(A.shape[0]), 0)
    X = np.random.default_rng(0).standard_normal((A.shape[0], 2))
    HX = (B @ X).cpu().numpy()
    st = oscf.build_setup(atoms, dom, "cube", 8)
    ref = (0.5 * st.A + st.vnl) @ X
    assert np.allclose(HX, ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())
