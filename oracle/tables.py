"""Element table and pseudopotential spline tables (oracle side).

Data only: ``parsec_python_amd/data/elements.csv`` is the reference's
``src/elements_new.csv`` with a header row added, and
``parsec_python_amd/data/spline_tables.npz`` holds the 18 ``data<El>`` arrays of
the reference's ``src/V_ion/splineData.mat`` (columns radius, charge, hartree,
pot_P, pot_S, wfn_P, wfn_S -- ``src/V_ion/splineData.py:35``).
Column meanings of the element table follow ``src/rsdft_input.py:28-30``.
"""
from __future__ import annotations

import csv
import functools
import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                     "parsec_python_amd", "data")

COL_RADIUS, COL_CHARGE, COL_HARTREE, COL_POT_P, COL_POT_S, COL_WFN_P, COL_WFN_S = range(7)


@functools.lru_cache(maxsize=None)
def elements() -> dict:
    """Return {symbol: dict(Z, Zvalue, h, r, R)} from the element table."""
    out = {}
    with open(os.path.join(_DATA, "elements.csv"), newline="") as fh:
        for row in csv.DictReader(fh):
            out[row["Element"]] = {
                "Z": float(row["Z"]),
                "Zvalue": float(row["Zvalue"]),
                "h": float(row["h"]),
                "r": float(row["r"]),
                "R": float(row["R"]),
            }
    return out


@functools.lru_cache(maxsize=None)
def spline_table(symbol: str) -> np.ndarray:
    with np.load(os.path.join(_DATA, "spline_tables.npz"), allow_pickle=False) as d:
        return np.array(d[symbol], dtype=np.float64)
