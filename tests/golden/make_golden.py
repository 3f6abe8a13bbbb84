"""Writes the regression fixtures of tests/golden from the FP64 CPU oracle.

cavity8_5steps.npz: 8^3 lid-driven cavity (Re 100, dt 0.005, seeded 1e-3
perturbation, tol 1e-12) after 5 steps; padded u1, u2, u3, p in the
reference layout. The reference ships no vectors for this path, so this
fixture anchors the oracle against unintended changes only; the oracle's
correctness is pinned by the closed-form tests in tests/test_oracle.py.
"""
import importlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
pkg = importlib.import_module("2020researchproject_amd")

if __name__ == "__main__":
    s = pkg.ibm3d.Solver(os.path.join(ROOT, "oracle", "libibm3d_oracle.so"))
    pkg.cases.setup_cavity(s, n=8, tol=1e-12)
    s.iterate(5)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "cavity8_5steps.npz"),
                        **{n: s.get(n) for n in ("U1", "U2", "U3", "P")})
    print("wrote cavity8_5steps.npz")
