"""Writes the regression fixtures tests/golden/*.npz.

The reference ships no stored vectors (its tests are property checks), so
these fixtures are restatement-generated: inputs and the CPU oracle's outputs
for small seeded cases, one per (quadrature family, collision, boundary)
combination.  They pin the oracle (and, through it, the GPU path) against
accidental change; they are not evidence of agreement with Julia by
themselves -- the property and analytic tests are.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from lbm_cases import bc_dicts, random_populations, shear_wave_fields  # noqa: E402
from oracle import clib, lbm_numpy as ln  # noqa: E402

CASES = {
    # name: (qid, nx, ny, collision, taus, steps, force, bc list name, init)
    "d2q9_srt_shear16": (9, 16, 16, 0, [0.8], 50, None, "periodic", "shear"),
    "d2q9_trt_couette": (9, 6, 10, 1, [0.7, 1.3], 40, None, "couette", "random"),
    "d2q9_mrt_poiseuille": (9, 3, 8, 2, [0.9, 0.75], 25, [1e-5, 0.0], "poiseuille", "random"),
    "d2q13_srt_cavity": (13, 8, 8, 0, [0.9], 20, None, "cavity", "random"),
    "d2q17_trt_partial": (17, 9, 7, 1, [0.8, 1.1], 20, [2e-6, -1e-6], "partial", "random"),
    "d2q21_srt_couette": (21, 5, 9, 0, [0.85], 15, None, "couette", "random"),
    "d2q37_srt_poiseuille": (37, 4, 10, 0, [0.9], 15, [1e-5, 0.0], "poiseuille", "random"),
    "d2q4_srt_periodic": (4, 6, 5, 0, [0.7], 30, None, "periodic", "random"),
    "d2q5_trt_periodic": (5, 6, 5, 1, [0.7, 0.9], 30, None, "periodic", "random"),
}


def initial(qid, nx, ny, init):
    if init == "shear":
        rho, ux, uy = shear_wave_fields(nx, 0.01)
        return np.asfortranarray(ln.equilibrium_field(ln.by_id(qid), rho, ux, uy))
    return random_populations(qid, nx, ny, seed=qid * 1000 + nx)


def make(name):
    qid, nx, ny, coll, taus, steps, force, bcname, init = CASES[name]
    f0 = initial(qid, nx, ny, init)
    S, C = f0.copy(order="F"), f0.copy(order="F")
    F = None if force is None else np.array(force)
    clib.simulate(qid, S, C, coll, taus, steps, force=F, bcs=bc_dicts(bcname, nx, ny))
    return dict(f0=f0, f_stream=S, f_collision=C)


if __name__ == "__main__":
    clib.build()
    for name in CASES:
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **make(name))
        print("wrote", name)
