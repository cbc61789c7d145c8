"""Writes cantilever_tet10_60x12x12_skyline.npz: the displacements of the Tet10 cantilever
box 300x60x60 mm (60x12x12 cells x 6 Tet10, 225,000 free DOF; E = 70000, nu = 0.33, clamped
at x = 0, traction (0, 5, -10) MPa on x = 300) solved by the oracle's FP64 skyline Cholesky
(the reference's direct solver, solve.rs:10-139, restated in oracle/maille_oracle.cpp).

Self-generated by this repository's oracle, not by the reference (which cannot run here). Its
profile holds 402,531,399 entries (3.2 GB), which takes minutes on one host core: too long for
a test run, so the solution is kept as a fixture. Run from the repository root:
    python tests/golden/make_cantilever_tet10_skyline.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from maille_amd import meshes  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "cantilever_tet10_60x12x12_skyline.npz")


def main():
    p = meshes.cantilever_tet10()
    t = time.time()
    ref = oracle.model_solve(p, iterative=False)
    assert ref["status"] == oracle.OK, ref["status"]
    t_direct = time.time() - t
    t = time.time()
    pcg = oracle.model_solve(p, iterative=True)
    t_pcg = time.time() - t
    u, up = ref["displacements"], pcg["displacements"]
    gap = float(np.abs(up - u).max() / np.abs(u).max())
    l2 = float(np.linalg.norm(up - u) / np.linalg.norm(u))
    np.savez_compressed(OUT, displacements=u, nfree=np.int64(p.nfree),
                        oracle_pcg_relative_gap=gap, oracle_pcg_rel_l2=l2,
                        oracle_pcg_iterations=np.int64(pcg["iterations"]))
    print(f"skyline {t_direct:.0f} s, oracle PCG {t_pcg:.0f} s ({pcg['iterations']} iterations): "
          f"relative gap {gap:.2e}, rel L2 {l2:.2e} -> {OUT}")


if __name__ == "__main__":
    main()
