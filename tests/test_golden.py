"""Regression fixtures (restatement-generated, see tests/golden/make_golden.py):
the oracle reproduces them bit for bit, and on a GPU the HIP path does too."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as mg  # noqa: E402
from lbm_cases import bc_dicts, to_lbmhip  # noqa: E402


def _load(name):
    with np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(mg.CASES))
def test_oracle_reproduces_fixture(oracle_lib, name):
    g = _load(name)
    assert np.array_equal(mg.initial(*mg.CASES[name][:3], mg.CASES[name][8]), g["f0"])
    out = mg.make(name)
    assert np.array_equal(out["f_stream"], g["f_stream"])
    assert np.array_equal(out["f_collision"], g["f_collision"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(mg.CASES))
def test_gpu_reproduces_fixture(name):
    import lbmhip as L
    qid, nx, ny, coll, taus, steps, force, bcname, _ = mg.CASES[name]
    g = _load(name)
    q = L.Quadratures[f"D2Q{qid}"]()
    F = None if force is None else np.array(force)
    cm = {0: lambda: L.SRT(taus[0], F), 1: lambda: L.TRT.fields(taus[0], taus[1], F),
          2: lambda: L.MRT(q, taus, force=F)}[coll]()
    model = L.LatticeBoltzmannModel.from_populations(q, np.asfortranarray(g["f0"]), cm,
                                                     to_lbmhip(bc_dicts(bcname, nx, ny)))
    L.simulate(model, range(steps))
    assert np.array_equal(model.f_stream.to_numpy(), g["f_stream"])
    assert np.array_equal(model.f_collision.to_numpy(), g["f_collision"])
