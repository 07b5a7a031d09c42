"""What the trained initialiser gains: on held-out seeded geometries (seeds the
training never saw) the SCF started from the network's density reaches
|dE| < 1e-6 Ha/atom in fewer iterations than from SAD, at the same converged
energy.  Also checks the start-density contract of the reference's ML path
(src/V_ion/pseudoDiag_ML4Den_poisson.py:392-540): non-negative, normalised to
the electron count, V_H1 - V_H0 solves the Poisson correction."""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HELDOUT_SEED = 777001


def _net():
    from parsec_python_amd.learned import UNet, default_checkpoint
    path = default_checkpoint()
    if not os.path.exists(path):
        pytest.skip("no trained checkpoint packaged")
    return UNet.load(path)


def test_learned_start_contract(gpu):
    import torch
    from parsec_python_amd import molecules as M
    from parsec_python_amd.learned import TRAIN_H_ANG, TRAIN_RADIUS_ANG, learned_start
    from parsec_python_amd.scf import SCFConfig, SCFProblem
    net = _net()
    atoms, dom, nev = M.prepare(M.alkane(4, np.random.default_rng(HELDOUT_SEED)), h_ang=TRAIN_H_ANG,
                                radius_ang=TRAIN_RADIUS_ANG)
    prob = SCFProblem(atoms, dom, nev, SCFConfig(domain="cube", pcg_maxit=1000))
    rho1, vh1 = learned_start(prob, net)
    r = rho1.cpu().numpy()
    assert r.min() >= 0.0
    assert abs(r.sum() - prob.nelec) < 1e-9 * prob.nelec
    # (-lap/h^2)(V_H1 - V_H0) = 4 pi/h^3 (rho1 - rho0) to the CG tolerance
    from oracle import fd as ofd
    g = prob.grid
    A = ofd.laplacian_masked(g.nx, g.ny, g.nz, 4, g.mask) * (1.0 / g.h ** 2)
    dv = (vh1 - prob.hpot0).cpu().numpy()
    rhs = 4 * math.pi / g.h ** 3 * (r - prob.rho0.cpu().numpy())
    assert np.linalg.norm(A @ dv - rhs) <= 1e-8 * np.linalg.norm(rhs)
    prob.close()


def test_learned_start_saves_iterations(gpu):
    from parsec_python_amd.learned import iterations_with, sample_geometry
    net = _net()
    rng = np.random.default_rng(HELDOUT_SEED)
    sad, lrn = [], []
    for kind in ("alkane", "alkane", "alkane", "si"):
        geo = sample_geometry(kind, rng)
        its0, e0, c0 = iterations_with(geo, None)
        its1, e1, c1 = iterations_with(geo, net)
        print(f"held-out {kind}: SAD {its0} its, learned {its1} its, |dE| {abs(e1 - e0):.2e} Ha/atom", flush=True)
        assert c0 and c1
        assert abs(e1 - e0) < 1e-5
        sad.append(its0)
        lrn.append(its1)
    assert np.mean(lrn) < np.mean(sad)
