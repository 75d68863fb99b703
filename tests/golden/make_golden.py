"""Writes the golden fixtures of this directory from the CPU oracle.

The reference ships no golden vectors for this path (its tests compute their oracles live,
against MFEM), so the fixtures are the oracle's outputs on small configurations, pinned by
``tests/test_oracle.py``; they guard the oracle and the device path against regressions.

    python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import fe_oracle as O  # noqa: E402

CASES = [
    # name, n, order, perturb, density, numbering, blocks
    ("hex2_p1_linear", 2, 1, 0.0, "linear", "lattice"),
    ("hex2_p2_nonlinear_perturbed", 2, 2, 0.2, "nonlinear", "lattice"),
    ("hex4_p2_linear", 4, 2, 0.0, "linear", "lattice"),
    ("hex3_p2_linear_mfem", 3, 2, 0.0, "linear", "mfem"),
]

BLOCKS = {"A": (2, 0), "C": (2, 1), "Wum": (0, 1), "Wmm": (1, 1), "Wuu": (0, 0)}


def make(name, n, order, perturb, density, numbering):
    P = O.Problem(n, order=order, perturb=perturb, density=density, numbering=numbering)
    out = {"edofs_u": P.Vu.edofs.astype(np.int32), "edofs_m": P.Vm.edofs.astype(np.int32),
           "u": P.u, "m": P.m, "p": P.p, "ess": P.ess.astype(np.int64),
           "residual": P.residual(2)}
    for key, (i, j) in BLOCKS.items():
        if key == "Wuu" and density == "linear":
            continue
        ip, ix, d = P.assembled(i, j)
        out[key + "_indptr"], out[key + "_indices"], out[key + "_data"] = ip, ix, d
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)


if __name__ == "__main__":
    for case in CASES:
        make(*case)
        print("wrote", case[0])
