"""What the trained initialiser gains, on held-out seeded geometries (seeds the
training never saw): the start density's L1 distance to the converged density
shrinks several-fold against SAD, and the SCF started from it converges to the
same energy.  The SCF iteration counts from both starts are printed; their
effect on time-to-converge is what bench.py measures on Si87H76.  Also checks
the start-density contract of the reference's ML path
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
    from parsec_python_amd import molecules as M
    from parsec_python_amd.learned import TRAIN_H_ANG, TRAIN_RADIUS_ANG, learned_start
    from parsec_python_amd.scf import SCFConfig, SCFProblem
    net = _net()
    atoms, dom, nev = M.prepare(M.alkane(4, np.random.default_rng(HELDOUT_SEED)), h_ang=TRAIN_H_ANG,
                                radius_ang=TRAIN_RADIUS_ANG)
    prob = SCFProblem(atoms, dom, nev, SCFConfig(domain="cube", pcg_maxit=1000))
    from oracle import fd as ofd
    g = prob.grid
    A = ofd.laplacian_masked(g.nx, g.ny, g.nz, 4, g.mask) * (1.0 / g.h ** 2)
    # the reference's tolerance (pseudoDiag_ML4Den_poisson.py:532-538) by default, and a tight solve
    for tol, kw in ((1e-5, {}), (1e-10, {"cg_tol": 1e-10})):
        rho1, vh1 = learned_start(prob, net, **kw)
        r = rho1.cpu().numpy()
        assert r.min() >= 0.0
        assert abs(r.sum() - prob.nelec) < 1e-9 * prob.nelec
        # (-lap/h^2)(V_H1 - V_H0) = 4 pi/h^3 (rho1 - rho0) to the CG tolerance (zero start: r0 = rhs)
        dv = (vh1 - prob.hpot0).cpu().numpy()
        rhs = 4 * math.pi / g.h ** 3 * (r - prob.rho0.cpu().numpy())
        assert np.linalg.norm(A @ dv - rhs) <= 1.05 * tol * np.linalg.norm(rhs)
    prob.close()


def test_learned_start_is_closer_to_the_converged_density(gpu):
    from parsec_python_amd import molecules as M
    from parsec_python_amd.learned import TRAIN_H_ANG, TRAIN_RADIUS_ANG, learned_start, sample_geometry
    from parsec_python_amd.scf import SCFConfig, SCFProblem, run_scf
    net = _net()
    rng = np.random.default_rng(HELDOUT_SEED)
    for kind in ("alkane", "alkane", "alkane", "si"):
        atoms, dom, nev = M.prepare(sample_geometry(kind, rng), h_ang=TRAIN_H_ANG, radius_ang=TRAIN_RADIUS_ANG)
        cfg = SCFConfig(tol=1e-12, maxits=60, domain="cube", precision="mixed", etol_ha_per_atom=1e-6,
                        pcg_maxit=1000, pcg_aliasing=False, seed=1)
        prob = SCFProblem(atoms, dom, nev, cfg)
        r0 = run_scf(prob)
        rho_c = r0.rho * prob.h ** 3
        rho1, vh1 = learned_start(prob, net)
        r1 = run_scf(prob, rho_init=rho1, hpot_init=vh1)
        err_sad = float((prob.rho0 - rho_c).abs().sum()) / prob.nelec
        err_ml = float((rho1 - rho_c).abs().sum()) / prob.nelec
        e0 = 0.5 * r0.energies["total_ry"] / prob.n_atoms
        e1 = 0.5 * r1.energies["total_ry"] / prob.n_atoms
        print(f"held-out {kind}: start L1 error per electron SAD {err_sad:.3e} learned {err_ml:.3e}; "
              f"SCF its SAD {r0.iterations} learned {r1.iterations}; |dE| {abs(e1 - e0):.1e} Ha/atom", flush=True)
        assert r0.converged and r1.converged
        assert err_ml < 0.2 * err_sad
        assert abs(e1 - e0) < 1e-5
        prob.close()


def test_learned_start_beats_the_sad_weighted_control(gpu, tmp_path):
    """The gain of the learned start over its control: the SAD start with
    first_filt's start vectors weighted by the SAD density (the weighting the
    learned start also uses).  On held-out seeded geometries (4 geometries x 2
    start-vector seeds, deterministic) the learned start needs fewer SCF
    iterations on average, by at least one, and every run converges to the
    same energy."""
    from parsec_python_amd.learned import evaluate_main
    _net()
    rows, mean = evaluate_main(["--geometries", "4", "--seeds", "2", "--heldout-seed", str(HELDOUT_SEED),
                                "--out", str(tmp_path / "report.md")])
    print(f"mean SCF iterations {mean}", flush=True)
    assert mean["learned"] + 1.0 <= mean["sad_weighted"]
    assert mean["learned"] < mean["sad"]
    assert all(r["de"] < 1e-5 for r in rows)
    assert all(r["l1"][1] < 0.2 * r["l1"][0] for r in rows)
