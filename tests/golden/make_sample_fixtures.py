"""Copy the reference's sample run files (inputs and logs: data, not code)
and reduce its saved density arrays to small fixtures.

Reads /root/reference/samples (run here, never on a GPU machine):

* samples/h/h_smoke.json, samples/h2o/H2O_sad_diagmeth3_5A_0p2A.in,
  samples/h8c10/cpu/H8C10_sad_diagmeth3_9A_0p2A.in -> samples/ (inputs);
* samples/h/h_smoke.out, samples/h2o/H2O_sad_diagmeth3_5A_0p2A.out -> samples/
  (logs, for the .out reader and writer);
* the H2O init/conv density files in their three units (e/cell, e/bohr^3,
  e/Angstrom^3) -> density_marginals.npz: for each file its shape, total, and
  the sums over two of the three axes (one vector per axis).  The marginals
  fix the array orientation and the unit factor of each variant without
  committing the arrays.
"""
import os
import shutil

import numpy as np

REF = "/root/reference/samples"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "samples")

COPIES = ["h/h_smoke.json", "h/h_smoke.out", "h2o/H2O_sad_diagmeth3_5A_0p2A.in",
          "h2o/H2O_sad_diagmeth3_5A_0p2A.out", "h8c10/cpu/H8C10_sad_diagmeth3_9A_0p2A.in"]


def main():
    os.makedirs(OUT, exist_ok=True)
    for rel in COPIES:
        shutil.copyfile(os.path.join(REF, rel), os.path.join(OUT, os.path.basename(rel)))
    marg = {}
    for stage in ("init", "conv"):
        for suffix in ("_grid", "_bohr3", ""):
            name = f"H2O_sad_diagmeth3_5A_0p2A_{stage}_rho{suffix}.npy"
            a = np.load(os.path.join(REF, "h2o", name), allow_pickle=False)
            key = f"{stage}{suffix or '_ang3'}"
            marg[f"{key}_shape"] = np.array(a.shape)
            marg[f"{key}_total"] = np.array(a.sum())
            marg[f"{key}_x"] = a.sum(axis=(1, 2))
            marg[f"{key}_y"] = a.sum(axis=(0, 2))
            marg[f"{key}_z"] = a.sum(axis=(0, 1))
    np.savez_compressed(os.path.join(HERE, "density_marginals.npz"), **marg)


if __name__ == "__main__":
    main()
