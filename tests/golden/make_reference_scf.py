"""Write tight-tolerance reference vectors by running the reference's own CPU path.

Run in the development container only (it reads the reference checkout at
/root/reference, which does not exist where the GPU tests run):

    python tests/golden/make_reference_scf.py            # all cases
    python tests/golden/make_reference_scf.py --case h2o_baseline_cube

What no closed form or logged number pins, and this script records:

* the end-to-end FP64 SCF fixed point at a tight tolerance (the reference's
  logs stop at tol 2e-4, about 1e-4 Ry): total energy and its components,
  eigenvalues and occupations, for
    - ``h2o_sample_cube``: the reference's own H2O sample input
      (samples/h2o/H2O_sad_diagmeth3_5A_0p2A.in) with tol tightened;
    - ``h2o_baseline_cube``: H2O with O at the origin and H at
      (+-0.757, 0.587, 0) A, radius 5 A, h 0.2 A, nev 8;
  driven through the reference's own pipeline (prepare_input_data ->
  prepare_system -> select_solver_backend(use_gpu=0) ->
  run_rsdft_calculation, src/main_new.py:46-111), so the energy is the
  reference's compute_energy_components (src/rsdft_solver.py:128-153) at full
  double precision instead of the 5 printed decimals;
* the local pseudopotential column V_loc of pseudoDiag
  (src/V_ion/pseudoDiag.py:44-421) and the nonlocal operator V_nl of
  pseudoNL_original (src/V_ion/pseudoNL_original.py:9-208) applied to a seeded
  block, on a small H2O grid (h 0.5 A, radius 3 A).

Outputs (small fixtures, data only): tests/golden/reference_scf_tight.json and
tests/golden/reference_ionic_small.npz.  The SCF's random start vectors come
from numpy's global generator, seeded here; the fixed point does not depend
on them.
"""
from __future__ import annotations

import argparse
import contextlib
import io
import json
import os
import subprocess
import sys
import tempfile
import time

REF_SRC = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT_JSON = os.path.join(HERE, "reference_scf_tight.json")
OUT_NPZ = os.path.join(HERE, "reference_ionic_small.npz")

SAMPLE_SYSTEM = """O  -3.56626  1.77639  0.00000
H  -2.59626  1.77639  0.00000
H  -3.88959  1.36040 -0.81444"""
BASELINE_SYSTEM = """O   0.00000  0.00000  0.00000
H   0.75700  0.58700  0.00000
H  -0.75700  0.58700  0.00000"""

C10H8_SAMPLE_SYSTEM = "\n".join(
    ["H 9.000000 11.481313 10.220927", "H 9.000000 10.240247 12.355972", "H 9.000000 7.759753 12.355972",
     "H 9.000000 6.518687 10.220927", "H 9.000000 11.481313 7.779073", "H 9.000000 6.518687 7.779073",
     "H 9.000000 7.759753 5.644028", "H 9.000000 10.240247 5.644028", "C 9.000000 10.387757 10.228803",
     "C 9.000000 9.700099 11.407084", "C 9.000000 8.299901 11.407084", "C 9.000000 7.612243 10.228803",
     "C 9.000000 8.289915 9.000000", "C 9.000000 9.710085 9.000000", "C 9.000000 10.387757 7.771197",
     "C 9.000000 7.612243 7.771197", "C 9.000000 8.299901 6.592916", "C 9.000000 9.700099 6.592916"])

TIGHT_TOL = 1e-9

CASES = {
    "h2o_sample_cube": dict(system=SAMPLE_SYSTEM, h=0.2, radius=5.0, nev=None, tol=TIGHT_TOL, maxits=80),
    "h2o_baseline_cube": dict(system=BASELINE_SYSTEM, h=0.2, radius=5.0, nev=8, tol=TIGHT_TOL, maxits=80),
    # samples/h8c10/cpu/H8C10_sad_diagmeth3_9A_0p2A.in with the tolerance tightened (778,688 points, nev 34)
    "c10h8_sample_cube": dict(system=C10H8_SAMPLE_SYSTEM, h=0.2, radius=9.0, nev=None, tol=1e-8, maxits=60),
}


def _input_text(c: dict) -> str:
    lines = ["$system", c["system"], "$system", "", "$settings", "unit = Angstrom", "Z_charge = 0",
             "density_method = sad", f"grid_spacing = {c['h']:.6f}", f"sphere_radius = {c['radius']:.6f}",
             f"tol = {c['tol']:.6e}", "Fermi_temp = 500", f"maxits = {c['maxits']}", "fd_order = 8", "poldeg = 10",
             "diagmeth = 3", "save_wfn = 0", "use_gpu = 0", "recenter_atoms = 1"]
    if c.get("nev"):
        lines.append(f"nev = {c['nev']}")
    lines.append("$settings")
    return "\n".join(lines) + "\n"


def _import_reference():
    """Put the reference's src/ on the path and move to a scratch directory
    first (importing pseudoDiag opens a log file in the working directory)."""
    if not os.path.isdir(REF_SRC):
        raise SystemExit(f"{REF_SRC} not found: this script runs only where the reference checkout is present")
    os.chdir(tempfile.mkdtemp(prefix="refimport_"))
    if REF_SRC not in sys.path:
        sys.path.insert(0, REF_SRC)


def run_scf_case(name: str) -> dict:
    """One reference SCF run (in this process); returns the fixture record."""
    import numpy as np
    _import_reference()
    from rsdft_backend import select_solver_backend
    from rsdft_input import load_elements, prepare_input_data
    from rsdft_models import SolverSettings
    from rsdft_setup import prepare_system
    from rsdft_solver import compute_energy_components, run_rsdft_calculation

    c = CASES[name]
    work = tempfile.mkdtemp(prefix=f"refscf_{name}_")
    inp = os.path.join(work, f"{name}.in")
    with open(inp, "w") as f:
        f.write(_input_text(c))
    np.random.seed(20261203)
    settings = SolverSettings()
    settings.use_gpu = 0
    settings.normalize()
    elem, n_elements = load_elements(REF_SRC)
    log = io.StringIO()
    t0 = time.time()
    cwd = os.getcwd()
    os.chdir(work)
    try:
        with contextlib.redirect_stdout(log):
            input_data = prepare_input_data(elem, settings, inp, REF_SRC)
            settings.apply_overrides(input_data.settings_overrides)
            settings.use_gpu = 0
            settings.normalize()
            backend = select_solver_backend(settings)
            problem = prepare_system(input_data, settings, elem, n_elements)
            res = run_rsdft_calculation(problem, elem, n_elements, backend)
    finally:
        os.chdir(cwd)
    wall = time.time() - t0
    h = problem.domain["h"]
    e = compute_energy_components(res.rho, res.hpot, res.hpot0, res.xc_potential, res.exc, res.eigenvalues,
                                  res.occupations, problem.nev, res.e_nuc0, h)
    energies = {k: float(v) for k, v in vars(e).items()} if hasattr(e, "__dict__") else {
        k: float(getattr(e, k)) for k in e.__dataclass_fields__}
    rec = {
        "input": _input_text(c),
        "domain": {k: float(v) if isinstance(v, float) else v for k, v in problem.domain.items()},
        "tol": c["tol"],
        "nev": int(problem.nev),
        "n_atoms": int(res.n_atoms),
        "iterations": int(res.iterations),
        "final_error": float(res.error),
        "converged": bool(res.converged),
        "energies_ry": energies,
        "eigenvalues_ha": [float(x) for x in np.asarray(res.eigenvalues).reshape(-1)],
        "occupations": [float(x) for x in np.asarray(res.occupations).reshape(-1)],
        "rho_sum_e": float(np.sum(res.rho) * h ** 3),
        "wall_s": round(wall, 1),
        "atoms_bohr": [{"typ": a["typ"], "coord": np.asarray(a["coord"]).tolist()} for a in problem.input_data.atoms],
    }
    return rec


def run_ionic_small() -> dict:
    """V_loc (pseudoDiag) and V_nl @ X (pseudoNL_original) on a small H2O grid."""
    import numpy as np
    _import_reference()
    from rsdft_input import load_elements, prepare_input_data
    from rsdft_models import SolverSettings
    from rsdft_setup import prepare_system
    from V_ion.pseudoDiag import pseudoDiag
    from V_ion.pseudoNL_original import pseudoNL

    c = dict(system=BASELINE_SYSTEM, h=0.5, radius=3.0, nev=4, tol=1e-4, maxits=1)
    work = tempfile.mkdtemp(prefix="refionic_")
    inp = os.path.join(work, "h2o_small.in")
    with open(inp, "w") as f:
        f.write(_input_text(c))
    settings = SolverSettings()
    settings.normalize()
    elem, n_elements = load_elements(REF_SRC)
    log = io.StringIO()
    cwd = os.getcwd()
    os.chdir(work)
    try:
        with contextlib.redirect_stdout(log):
            input_data = prepare_input_data(elem, settings, inp, REF_SRC)
            settings.apply_overrides(input_data.settings_overrides)
            settings.normalize()
            problem = prepare_system(input_data, settings, elem, n_elements)
            out = pseudoDiag(problem.domain, problem.input_data.atoms, elem, n_elements, return_info=True)
            rho0, hpot0, ppot = out[0], out[1], out[2]
            vnl = pseudoNL(problem.domain, problem.input_data.atoms, elem, n_elements)
    finally:
        os.chdir(cwd)
    n = problem.domain["nx"] * problem.domain["ny"] * problem.domain["nz"]
    rng = np.random.default_rng(7)
    X = rng.standard_normal((n, 3))
    VX = vnl @ X
    coords = np.concatenate([np.asarray(a["coord"], dtype=np.float64) for a in problem.input_data.atoms])
    typs = np.array(sum(([a["typ"]] * len(a["coord"]) for a in problem.input_data.atoms), []))
    return {"nx": np.int64(problem.domain["nx"]), "h": np.float64(problem.domain["h"]),
            "radius": np.float64(problem.domain["radius"]), "coords_bohr": coords, "species": typs,
            "rho0": np.asarray(rho0, dtype=np.float64).reshape(-1), "hpot0": np.asarray(hpot0).reshape(-1),
            "ppot": np.asarray(ppot).reshape(-1), "X": X, "VnlX": np.asarray(VX)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", choices=sorted(CASES) + ["ionic_small"])
    ap.add_argument("--json-out", help="write one case's record here (used by the parent run)")
    ap.add_argument("--add", nargs="+", choices=sorted(CASES), help="run these cases and add them to the fixture")
    args = ap.parse_args()
    if args.case == "ionic_small":
        import numpy as np
        np.savez_compressed(OUT_NPZ, **run_ionic_small())
        print(f"wrote {OUT_NPZ}")
        return
    if args.case:
        rec = run_scf_case(args.case)
        with open(args.json_out or os.path.join(tempfile.gettempdir(), f"{args.case}.json"), "w") as f:
            json.dump(rec, f)
        print(f"{args.case}: {rec['iterations']} iterations, err {rec['final_error']:.2e}, "
              f"E_total {rec['energies_ry']['total_ry']:.10f} Ry, {rec['wall_s']} s")
        return
    # every case in its own process (the reference keeps mixer state in module globals)
    out = {"generator": "tests/golden/make_reference_scf.py", "cases": {}}
    names = list(CASES)
    if args.add:
        with open(OUT_JSON) as f:
            out = json.load(f)
        names = args.add
    for name in names:
        tmp = os.path.join(tempfile.gettempdir(), f"refscf_{name}.json")
        subprocess.run([sys.executable, __file__, "--case", name, "--json-out", tmp], check=True)
        with open(tmp) as f:
            out["cases"][name] = json.load(f)
    with open(OUT_JSON, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {OUT_JSON}")
    if not args.add:
        subprocess.run([sys.executable, __file__, "--case", "ionic_small"], check=True)


if __name__ == "__main__":
    main()
