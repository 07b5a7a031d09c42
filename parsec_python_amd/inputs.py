"""Run input files: the reference's ``.in``/``.inp``, ``.json``, ``.txt``,
``.dat`` and ``.mat`` geometry/settings formats.

Mirrors ``load_atomic_data_from_file`` (``src/rsdft_input.py:331-468``),
``convert_units_to_bohr`` (``src/rsdft_input.py:582-632``) and the settings
keys of ``SolverSettings.apply_overrides`` (``src/rsdft_models.py:18-124``):

* ``.in``/``.inp``: ``$system`` ... ``$system`` holds ``TYPE x y z`` lines,
  ``$settings`` ... ``$settings`` holds ``key = value`` lines; ``Z_charge``
  (or ``charge``), ``density_method`` (or ``density``), ``ml_file_path`` (or
  ``density_file``) and ``unit`` are taken out, every other key is a solver
  override;
* ``.json``: ``atoms`` as a list of ``{typ, coord}`` entries or of
  ``"TYPE x y z"`` strings (or one multi-line string), ``Z_charge``/``charge``,
  ``density_method``, ``ml_file_path``/``density_file``, ``settings`` (or
  ``params``) and top-level ``nev``, ``grid_spacing``, ``h``,
  ``sphere_radius``, ``radius``, ``recenter_atoms``;
* ``.txt``: ``TYPE x y z`` lines; ``.dat``: species count, then per species a
  type line and a line of coordinates; ``.mat``: an ``AtomsInMolecule``
  struct array.

Coordinates from a file are in Angstrom unless ``unit`` says Bohr; grid
spacing and radius overrides are converted with them.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from .tables import A0_ANG

_ANG = ("ang", "angstrom", "a")
_BOHR = ("bohr", "au", "a.u.")


@dataclass
class RunInput:
    """A parsed input file: species blocks in file order, raw units."""
    atoms: list                     # [(symbol, (n, 3) array)]
    z_charge: float = 0.0
    density_method: str = "sad"
    ml_file_path: str | None = None
    settings: dict[str, Any] = field(default_factory=dict)
    unit: str | None = None
    path: str | None = None

    @property
    def n_atom(self) -> list[int]:
        return [int(c.shape[0]) for _, c in self.atoms]


def _atoms_from_lines(lines) -> list:
    """``TYPE x y z`` lines -> species blocks in order of first appearance
    (src/rsdft_input.py:198-226)."""
    grouped: dict[str, list] = {}
    for raw in lines:
        line = str(raw).strip()
        if not line or line.startswith("#"):
            continue
        parts = line.split()
        if len(parts) != 4:
            raise ValueError("Atom line must be: TYPE x y z")
        try:
            xyz = [float(p) for p in parts[1:]]
        except ValueError as exc:
            raise ValueError("Could not parse coordinates; must be numeric.") from exc
        grouped.setdefault(parts[0], []).append(xyz)
    return [(s, np.asarray(c, dtype=float)) for s, c in grouped.items()]


def _prune(d: dict) -> dict:
    return {k: v for k, v in (d or {}).items() if v is not None and not (isinstance(v, str) and v.strip() == "")}


def _resolve_relative(p: str | None, base: str) -> str | None:
    if not p:
        return None
    q = os.path.expanduser(str(p))
    return q if os.path.isabs(q) else os.path.join(os.path.dirname(base), q)


def read_input(path: str) -> RunInput:
    ext = os.path.splitext(path)[1].lower()
    if len(os.path.basename(path)) <= 4:
        raise ValueError("File name is too short, must include file extension.")
    z_charge, method, ml, unit = 0.0, "sad", None, None
    settings: dict[str, Any] = {}
    if ext in (".in", ".inp"):
        section, atom_lines, set_lines = None, [], []
        with open(path, encoding="utf-8") as f:
            for raw in f:
                line = raw.strip()
                if not line or line.startswith("#"):
                    continue
                tok = line.lower()
                if tok in ("$system", "$settings"):
                    name = tok[1:]
                    section = name if section != name else None
                    continue
                if section == "system":
                    atom_lines.append(line)
                elif section == "settings":
                    set_lines.append(line)
        atoms = _atoms_from_lines(atom_lines)
        for entry in set_lines:
            if "=" not in entry:
                continue
            key, val = (s.strip() for s in entry.split("=", 1))
            if not val:
                continue
            kl = key.lower()
            if kl in ("z_charge", "charge"):
                try:
                    z_charge = float(val)
                except ValueError:
                    z_charge = 0.0
            elif kl in ("density_method", "density"):
                method = val.lower()
            elif kl in ("ml_file_path", "density_file"):
                ml = val
            elif kl == "unit":
                unit = val
            else:
                settings[key] = val
        ml = _resolve_relative(ml, path)
    elif ext == ".json":
        with open(path, encoding="utf-8") as f:
            data = json.load(f)
        block = data.get("atoms") or data.get("Atoms")
        if not block:
            raise ValueError('JSON file missing "atoms" list.')
        if isinstance(block, str):
            atoms = _atoms_from_lines(block.strip().splitlines())
        elif isinstance(block, list) and isinstance(block[0], str):
            atoms = _atoms_from_lines(block)
        else:
            atoms = []
            for e in block:
                typ = e.get("typ") or e.get("type") or e.get("element")
                xyz = e.get("coord") or e.get("coords") or e.get("coordinates")
                if typ is None or xyz is None:
                    raise ValueError("Each atom entry must have typ and coord/coords.")
                atoms.append((str(typ), np.asarray(xyz, dtype=float).reshape(-1, 3)))
        try:
            z_charge = float(data.get("Z_charge", data.get("charge", 0.0)))
        except (TypeError, ValueError):
            z_charge = 0.0
        method = data.get("density_method", "sad")
        method = method.lower() if isinstance(method, str) else "sad"
        ml = _resolve_relative(data.get("ml_file_path") or data.get("density_file"), path)
        settings = data.get("settings", data.get("params", {})) or {}
        settings = dict(settings) if isinstance(settings, dict) else {}
        unit = settings.pop("unit", None)
        for key in ("nev", "grid_spacing", "h", "sphere_radius", "radius", "grid_npy_path", "grid_poscar_path",
                    "recenter_atoms"):
            if key in data and key not in settings:
                settings[key] = data[key]
    elif ext == ".txt":
        with open(path, encoding="utf-8") as f:
            atoms = _atoms_from_lines(f.readlines())
    elif ext == ".dat":
        atoms = []
        with open(path, encoding="utf-8") as f:
            for _ in range(int(f.readline().strip())):
                typ = f.readline().strip()
                xyz = np.array([float(x) for x in f.readline().split()], dtype=float)
                if xyz.size % 3:
                    raise ValueError("Coordinate count in .dat must be divisible by 3.")
                atoms.append((typ, xyz.reshape(-1, 3)))
    elif ext == ".mat":
        from scipy.io import loadmat
        mat = loadmat(path)
        if "AtomsInMolecule" not in mat:
            raise ValueError("File does not contain the correct variable: AtomsInMolecule")
        atoms = []
        for e in np.asarray(mat["AtomsInMolecule"]).ravel():
            typ, xyz = e["typ"], e["coord"]
            typ = str(np.asarray(typ).ravel()[0]) if not isinstance(typ, str) else typ
            atoms.append((typ, np.asarray(xyz, dtype=float).reshape(-1, 3)))
    else:
        raise ValueError("File extension not recognized (supported: .dat, .mat, .json, .txt, .in, .inp).")
    if method not in ("sad", "ml", "sad_ml_grid"):
        method = "sad"
    return RunInput(atoms=atoms, z_charge=z_charge, density_method=method, ml_file_path=ml,
                    settings=_prune(settings), unit=unit, path=os.path.abspath(path))


def to_bohr(inp: RunInput):
    """(atoms in bohr, overrides with lengths in bohr): files default to
    Angstrom (src/rsdft_input.py:582-632)."""
    unit = (inp.unit or "ang").lower()
    if unit in _BOHR:
        factor = 1.0
    else:
        factor = 1.0 / A0_ANG
    atoms = [(s, np.asarray(c, dtype=float) * factor) for s, c in inp.atoms]
    over = dict(inp.settings)
    for key in ("grid_spacing", "h", "sphere_radius", "radius"):
        if key in over:
            try:
                over[key] = float(over[key]) * factor
            except (TypeError, ValueError):
                pass
    return atoms, over


def _bool(v) -> int:
    if isinstance(v, str):
        return 1 if v.strip().lower() in ("1", "true", "yes", "y", "on") else 0
    return 1 if v else 0


def _tol(v) -> float:
    """A tolerance given as an integer n >= 1 means 10^-n (rsdft_models.py:38-47)."""
    t = float(v)
    if t >= 1 and abs(t - round(t)) < 1e-12:
        return 10.0 ** (-int(round(t)))
    return t


def solver_settings(overrides: dict) -> dict:
    """Reference solver settings (defaults of src/rsdft_models.py:26-36) with
    the overrides applied; keys are matched case-insensitively."""
    s = {"fd_order": 8, "maxits": 40, "tol": 2.0e-4, "fermi_temp": 500.0, "save_wfn": 0, "cg_prec": 0,
         "poldeg": 10, "diagmeth": 3, "adaptive_scheme": 0, "use_gpu": 0, "recenter_atoms": 1, "nev": None,
         "h": None, "radius": None, "mixer": "msecant1"}
    alias = {"fd_order": "fd_order", "maxits": "maxits", "tol": "tol", "fermi_temp": "fermi_temp",
             "save_wfn": "save_wfn", "write_wfn": "save_wfn", "cg_prec": "cg_prec", "poldeg": "poldeg",
             "diagmeth": "diagmeth", "gpu": "use_gpu", "recenter": "recenter_atoms",
             "adaptivescheme": "adaptive_scheme", "adaptive_scheme": "adaptive_scheme", "use_gpu": "use_gpu",
             "recenter_atoms": "recenter_atoms", "nev": "nev", "grid_spacing": "h", "h": "h",
             "sphere_radius": "radius", "radius": "radius", "mixer": "mixer"}
    for key, val in overrides.items():
        k = alias.get(str(key).lower())
        if k is None:
            continue
        if k == "tol":
            s[k] = _tol(val)
        elif k in ("fermi_temp", "h", "radius"):
            s[k] = float(val)
        elif k in ("save_wfn", "cg_prec", "adaptive_scheme", "use_gpu", "recenter_atoms"):
            s[k] = _bool(val)
        elif k == "mixer":
            s[k] = str(val).lower()
        else:
            s[k] = int(float(val))
    s["maxits"] = max(1, s["maxits"])
    s["poldeg"] = max(1, s["poldeg"])
    if not 0 <= s["diagmeth"] <= 3:
        s["diagmeth"] = 3
    return s
