"""End-to-end GPU SCF against the FP64 oracle on the same inputs.

Both sides draw their random start vectors from the same counter-based stream
(tests/rng_mirror.py), so the FP64 trajectories agree to rounding.  The mixed
precision path is checked against the FP64 GPU path at tight SCF tolerance
(energy within 1e-6 Ha/atom, occupied eigenvalues within 1e-5 Ha).
"""
import numpy as np
import pytest

from oracle import scf as oscf
from tests.rng_mirror import MirrorRNG

pytestmark = pytest.mark.gpu


def _small_h2o():
    from parsec_python_amd import molecules as M
    return M.prepare(M.h2o(), h_ang=0.3, radius_ang=4.0, nev=8)


@pytest.mark.parametrize("kind", ["cube", "sphere"])
def test_scf_fp64_matches_oracle_trajectory(gpu, kind):
    from parsec_python_amd.scf import SCFConfig, SCFProblem, run_scf
    atoms, dom, nev = _small_h2o()
    print(f"scf parity {kind} nx={dom['nx']}", flush=True)
    seed = 11
    cfg = SCFConfig(tol=1e-6, maxits=30, domain=kind, seed=seed)
    prob = SCFProblem(atoms, dom, nev, cfg)
    res = run_scf(prob)
    st = oscf.build_setup(atoms, dom, kind, 8)
    ref = oscf.run_scf(st, nev, tol=1e-6, maxits=30, poldeg=10, rng=MirrorRNG(seed))
    assert res.iterations == ref.iterations
    for hg, hr in zip(res.history, ref.history):
        assert hg["total_ry"] == pytest.approx(hr["total_ry"], abs=1e-8)
        assert np.allclose(hg["lam"], hr["lam"], atol=1e-9)
        assert hg["cg_its"] == hr["cg_its"]
    assert res.energies["total_ry"] == pytest.approx(ref.energies["total_ry"], abs=1e-9)
    assert np.allclose(res.rho.cpu().numpy(), ref.rho, rtol=1e-7, atol=1e-10)
    prob.close()


def test_scf_mixed_precision_matches_fp64(gpu):
    from parsec_python_amd.scf import SCFConfig, SCFProblem, run_scf
    atoms, dom, nev = _small_h2o()
    out = {}
    for prec in ("fp64", "mixed"):
        cfg = SCFConfig(tol=1e-8, maxits=40, domain="sphere", seed=3, precision=prec)
        prob = SCFProblem(atoms, dom, nev, cfg)
        out[prec] = run_scf(prob)
        prob.close()
    e64 = 0.5 * out["fp64"].energies["total_ry"] / 3
    e32 = 0.5 * out["mixed"].energies["total_ry"] / 3
    assert abs(e64 - e32) < 1e-6
    occ = out["fp64"].occupations > 0.5
    assert np.max(np.abs(out["fp64"].eigenvalues[:nev][occ] - out["mixed"].eigenvalues[:nev][occ])) < 1e-5


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_scf_overlapped_rayleigh_ritz_gives_the_same_bits(gpu, precision):
    """S = Y^T Y and its Cholesky factor on the dense stage's second stream
    (while H Y and G are formed on the main one) against the same steps in
    sequence on one stream: the same SCF to the last bit."""
    import torch
    from parsec_python_amd.scf import SCFConfig, SCFProblem, run_scf
    atoms, dom, nev = _small_h2o()
    out = {}
    for overlap in (False, True):
        cfg = SCFConfig(tol=1e-7, maxits=40, domain="sphere", seed=3, precision=precision, rr_overlap=overlap)
        prob = SCFProblem(atoms, dom, nev, cfg)
        res = run_scf(prob)
        out[overlap] = (res.iterations, res.energies["total_ry"], res.rho.clone(), np.array(res.eigenvalues))
        prob.close()
    print(f"  {precision}: iterations {out[False][0]} / {out[True][0]}", flush=True)
    assert out[False][0] == out[True][0]
    assert out[False][1] == out[True][1]
    assert torch.equal(out[False][2], out[True][2])
    assert np.array_equal(out[False][3], out[True][3])
