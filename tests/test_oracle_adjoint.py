"""Pins the reverse-mode checker (oracle/lbm_adjoint_numpy.py) on the CPU.

The reference has no gradient code; its gradients are those of its forward
run.  So the checker is pinned by:

* central finite differences of the C restatement's FP64 forward run
  (collide -> stream, lattice_boltzmann_model.jl:60-70) along random
  directions in u0 and along tau (tau_s, tau_a for TRT), every quadrature;
* the closed form at T = 0: u(feq(rho0, u0)) = u0, so dL/du0 = u0 - u_target;
* agreement of its own forward loss with the C restatement's.
"""
import numpy as np
import pytest

from lbm_cases import QIDS
from oracle import lbm_adjoint_numpy as la
from oracle import lbm_numpy as ln


def _fields(nx, ny, seed, amp=0.03):
    rng = np.random.default_rng(seed)
    u0 = amp * rng.standard_normal((nx, ny, 2))
    ut = amp * rng.standard_normal((nx, ny, 2))
    return u0, ut


def _c_loss(oracle_lib, qid, u0, ut, coll, taus, nsteps, rho0=1.0):
    """L from the C restatement: per-node equilibrium start, simulate, velocity."""
    nx, ny = u0.shape[:2]
    Q = ln.by_id(qid).Q
    f = np.empty((nx, ny, Q), order="F")
    for x in range(nx):
        for y in range(ny):
            f[x, y] = oracle_lib.equilibrium(qid, rho0, u0[x, y])
    S, C = f.copy(order="F"), f.copy(order="F")
    oracle_lib.simulate(qid, S, C, coll, taus, nsteps)
    _, ux, uy = oracle_lib.moments(qid, S, taus[0])[:3]
    return 0.5 * float(np.sum((ux - ut[..., 0]) ** 2 + (uy - ut[..., 1]) ** 2))


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("coll", [0, 1])
def test_checker_matches_central_differences(oracle_lib, qid, coll):
    q = ln.by_id(qid)
    nx, ny, T = 7, 6, 9
    taus = [0.8] if coll == 0 else [0.8, 1.3]
    u0, ut = _fields(nx, ny, seed=qid + 100 * coll)
    L, g, gt = la.gradient(q, u0, ut, taus, coll, T)
    assert abs(L / _c_loss(oracle_lib, qid, u0, ut, coll, taus, T) - 1.0) < 1e-12
    h = 1e-5
    rng = np.random.default_rng(7)
    for _ in range(2):
        v = rng.standard_normal(u0.shape)
        fd = (_c_loss(oracle_lib, qid, u0 + h * v, ut, coll, taus, T)
              - _c_loss(oracle_lib, qid, u0 - h * v, ut, coll, taus, T)) / (2 * h)
        ad = float(np.sum(g * v))
        assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
    for k in range(len(taus)):
        tp, tm = list(taus), list(taus)
        tp[k] += h
        tm[k] -= h
        fd = (_c_loss(oracle_lib, qid, u0, ut, coll, tp, T)
              - _c_loss(oracle_lib, qid, u0, ut, coll, tm, T)) / (2 * h)
        assert abs(gt[k] - fd) <= 1e-6 * abs(fd) + 1e-13, (k, gt[k], fd)


@pytest.mark.parametrize("qid", [9, 37])
def test_checker_zero_steps_closed_form(qid):
    q = ln.by_id(qid)
    u0, ut = _fields(5, 4, seed=3)
    L, g, gt = la.gradient(q, u0, ut, [0.7], 0, 0)
    np.testing.assert_allclose(g, u0 - ut, rtol=0, atol=1e-15)
    assert gt == [0.0]
    assert abs(L - 0.5 * np.sum((u0 - ut) ** 2)) < 1e-15


def test_checker_tau_gradient_sign_on_decay():
    """A decaying shear wave with u_target = 0: a larger tau (more viscosity)
    lowers the loss, dL/dtau < 0; a larger u0 raises it, <dL/du0, u0> > 0."""
    q = ln.by_id(9)
    n = 16
    y = np.arange(n)
    u0 = np.zeros((n, n, 2))
    u0[..., 0] = 0.01 * np.sin(2 * np.pi * y / n)[None, :]
    L, g, gt = la.gradient(q, u0, np.zeros_like(u0), [0.8], 0, 40)
    assert gt[0] < 0
    assert np.sum(g * u0) > 0
    # L is quadratic in the amplitude to leading order: <g, u0> ~ 2 L
    assert abs(np.sum(g * u0) / (2 * L) - 1.0) < 1e-3
