"""Pins the oracle's iterative initialization of Mei et al.
(initial_conditions/mei_et_al.jl:13-41, collision_models/iterative_initialization.jl).

* the nonlinear term carries no mass, momentum u and momentum flux u u
  (sum_i N_i = 0, sum_i c_i N_i = u, sum_i c_a c_b N_i = u_a u_b for
  quadratures with fourth-order isotropy), which fixes its s scalings;
* the collision with tau = 1 returns w rho + N exactly and conserves mass;
* a uniform prescribed velocity on a periodic lattice converges in two steps
  to w + N (density 1 at node (nx, ny) twice in a row);
* on a Taylor-Green velocity the converged state carries the prescribed
  velocity to the order of the Mach number squared;
* the C restatement against the NumPy one.
"""
import numpy as np
import pytest

from lbm_cases import bc_dicts, random_populations
from oracle import lbm_numpy as ln


def _tgv_u(nx, ny, u0=0.01):
    x = (np.arange(nx) + 0.5) * 2 * np.pi / nx
    y = (np.arange(ny) + 0.5) * 2 * np.pi / ny
    X, Y = np.meshgrid(x, y, indexing="ij")
    return np.asfortranarray(np.stack([u0 * np.cos(X) * np.sin(Y), -u0 * np.sin(X) * np.cos(Y)],
                                      axis=-1))


@pytest.mark.parametrize("qid", [9, 13, 17, 37])
def test_nonlinear_term_moments(oracle_lib, qid):
    q = ln.by_id(qid)
    u = _tgv_u(6, 5, 0.05)
    N = oracle_lib.mei_nonlinear(qid, u)
    c = q.c.astype(float)
    np.testing.assert_allclose(N.sum(axis=2), 0.0, atol=1e-17)
    np.testing.assert_allclose(np.einsum("xyi,di->xyd", N, c), u, rtol=1e-13, atol=1e-16)
    flux = np.einsum("xyi,ai,bi->xyab", N, c, c)
    np.testing.assert_allclose(flux, np.einsum("xya,xyb->xyab", u, u), rtol=1e-12, atol=1e-17)


@pytest.mark.parametrize("qid", [4, 5, 9, 21, 37])
def test_c_nonlinear_matches_numpy(oracle_lib, qid):
    q = ln.by_id(qid)
    u = _tgv_u(7, 4, 0.03)
    np.testing.assert_allclose(oracle_lib.mei_nonlinear(qid, u), ln.mei_nonlinear(q, u),
                               rtol=1e-13, atol=1e-19)


@pytest.mark.parametrize("qid", [9, 17])
@pytest.mark.parametrize("tau", [1.0, 0.7])
def test_collide_mei(oracle_lib, qid, tau):
    q = ln.by_id(qid)
    f = random_populations(qid, 6, 5, seed=qid)
    N = oracle_lib.mei_nonlinear(qid, _tgv_u(6, 5))
    out = np.zeros_like(f, order="F")
    oracle_lib.collide_mei(qid, f, out, tau, N)
    np.testing.assert_allclose(out, ln.collide_mei(q, f, tau, N), rtol=1e-14, atol=1e-17)
    np.testing.assert_allclose(out.sum(axis=2), f.sum(axis=2), rtol=1e-14)
    if tau == 1.0:
        rho = f[:, :, 0] + f[:, :, 1]
        for i in range(2, q.Q):
            rho = rho + f[:, :, i]
        assert np.array_equal(out, q.w * rho[..., None] + N)


@pytest.mark.parametrize("qid", [9, 37])
def test_uniform_velocity_converges_in_two_steps(oracle_lib, qid):
    q = ln.by_id(qid)
    nx, ny = 5, 4
    u = np.asfortranarray(np.broadcast_to([0.02, -0.01], (nx, ny, 2)).copy())
    fs, fc, steps = oracle_lib.mei_initialize(qid, u, tau=1.0, eps=1e-7)
    assert steps == 2
    N = oracle_lib.mei_nonlinear(qid, u)
    np.testing.assert_allclose(fs, q.w + N, rtol=1e-14, atol=1e-18)


def test_tgv_converges_to_the_prescribed_velocity(oracle_lib):
    qid, nx, ny, u0 = 9, 32, 32, 0.01
    q = ln.by_id(qid)
    u = _tgv_u(nx, ny, u0)
    fs, fc, steps = oracle_lib.mei_initialize(qid, u, tau=1.0, eps=1e-9, max_steps=10000)
    assert 2 < steps < 10000
    rho, ux, uy = ln.moments_field(q, fs)
    err = np.sqrt(np.sum((ux - u[..., 0]) ** 2 + (uy - u[..., 1]) ** 2) / np.sum(u ** 2))
    assert err < 5e-2, err
    # the density relaxed to a field of order u0^2 about 1
    assert np.abs(rho - 1).max() < 10 * u0 ** 2


def test_bounds_error_on_bad_bc(oracle_lib):
    u = _tgv_u(5, 4)
    bad = bc_dicts("poiseuille", 5, 4)
    bad[0]["xs"] = (1, 9)
    with pytest.raises(IndexError):
        oracle_lib.mei_initialize(9, u, bcs=bad)
