"""Element table and pseudopotential spline tables (package data).

``data/elements.csv`` is the reference's ``src/elements_new.csv`` with a header
(columns per ``src/rsdft_input.py:28-30``: Element, AtomicNumber, Z (valence),
Zvalue (KB denominator), h (default spacing, bohr), r (atom extent, bohr),
R (KB cutoff, bohr)).  ``data/spline_tables.npz`` holds the 18 ``data<El>``
tables of ``src/V_ion/splineData.mat`` with columns radius, charge, hartree,
pot_P, pot_S, wfn_P, wfn_S (``src/V_ion/splineData.py:35``).
"""
from __future__ import annotations

import csv
import functools
import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")

RADIUS, CHARGE, HARTREE, POT_P, POT_S, WFN_P, WFN_S = range(7)
A0_ANG = 0.529177210903
RY_EV = 13.605698066


@functools.lru_cache(maxsize=None)
def elements() -> dict:
    out = {}
    with open(os.path.join(_DATA, "elements.csv"), newline="") as fh:
        for row in csv.DictReader(fh):
            out[row["Element"]] = {k: float(row[k]) for k in ("Z", "Zvalue", "h", "r", "R")}
    return out


@functools.lru_cache(maxsize=None)
def table(symbol: str) -> np.ndarray:
    with np.load(os.path.join(_DATA, "spline_tables.npz"), allow_pickle=False) as d:
        if symbol not in d.files:
            raise KeyError(f"no pseudopotential table for {symbol!r}")
        return np.array(d[symbol], dtype=np.float64)
