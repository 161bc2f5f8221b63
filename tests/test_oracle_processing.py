"""Pins the oracle's processing sums: one row of TrackHydrodynamicErrors
(track_hydrodynamic_errors.jl:52-222) and the stopping-criteria sums
(stopping_criteria.jl:17-115).

Closed forms, each chosen so a plausible slip fails it:
* an equilibrium state of the expected (rho, u_max u) has zero density,
  velocity and pressure error, P = rho I, and sigma = 0 so each relative
  stress error is exactly 1;
* a shear perturbation eps w_i c_ix c_iy gives sigma_xy = eps cs^2 / (1 + 1/(2tau))
  divided by u_max^2 (dimensionless stress), with cs = 1/s;
* a trace perturbation eps w_i (c_ix^2 - cs) leaves rho and u alone and gives
  p = rho + eps cs^2 / (1 + 1/(2tau)) (a_2 from a_bar_2 and a_eq_2, line 153);
* a velocity off by a factor (1 + kappa) has error_u = |kappa|;
* uniform states give mass, momentum and energy in closed form;
* the affine factors are the convention offset + factor * plane;
* the C restatement against the NumPy one, every quadrature.
"""
import numpy as np
import pytest

from lbm_cases import QIDS, random_populations
from oracle import lbm_numpy as ln

ISO4 = [9, 13, 17, 37]   # quadratures with fourth-order isotropy and s as tabulated


def _smooth_fields(nx, ny, seed):
    rng = np.random.default_rng(seed)
    x = np.arange(nx)[:, None] / nx
    y = np.arange(ny)[None, :] / ny
    k = rng.uniform(0.5, 1.5, 4)
    rho = 1.0 + 0.01 * np.cos(2 * np.pi * (x + k[0] * y))
    ux = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * k[1] * y)
    uy = -np.cos(2 * np.pi * k[2] * x) * np.sin(2 * np.pi * y)
    return rho, ux, uy


def _planes(*fields):
    return np.asfortranarray(np.stack(fields, axis=-1))


IDENT = np.tile([0.0, 1.0], 8)


@pytest.mark.parametrize("qid", ISO4 + [21])
def test_equilibrium_state_has_zero_errors_and_unit_stress_error(oracle_lib, qid):
    q = ln.by_id(qid)
    nx, ny, u_max, tau = 13, 9, 0.02, 0.7
    rho, ux, uy = _smooth_fields(nx, ny, qid)
    f = np.asfortranarray(ln.equilibrium_field(q, rho, u_max * ux, u_max * uy))
    rng = np.random.default_rng(1)
    sig = 1.0 + rng.random((4, nx, ny))
    E = _planes(rho, ux, uy, rho, *sig)
    out = oracle_lib.track_errors(qid, f, E, IDENT, tau, u_max)
    assert out[0] < 1e-13
    assert out[1] < 1e-12
    assert out[2] < 1e-14
    np.testing.assert_allclose(out[3:7], 1.0, rtol=0, atol=1e-9)
    assert out[10] == 0


@pytest.mark.parametrize("qid", ISO4)
def test_shear_stress_closed_form(oracle_lib, qid):
    q = ln.by_id(qid)
    nx, ny, u_max, tau, eps = 5, 4, 0.05, 0.8, 1e-4
    cs = 1.0 / q.s
    f0 = ln.equilibrium_field(q, np.ones((nx, ny)), np.zeros((nx, ny)), np.zeros((nx, ny)))
    f = np.asfortranarray(f0 + eps * q.w * q.c[0] * q.c[1])
    sxy = eps * cs ** 2 / (1.0 + 1.0 / (2.0 * tau)) / u_max ** 2
    one = np.ones((nx, ny))
    E = _planes(one, one, one, one, one, sxy * one, sxy * one, one)
    out = oracle_lib.track_errors(qid, f, E, IDENT, tau, u_max)
    assert out[4] < 1e-11 and out[6] < 1e-11          # xy and yx match the closed form
    np.testing.assert_allclose([out[3], out[5]], 1.0, atol=1e-12)   # sigma_xx = sigma_yy = 0
    assert out[0] < 1e-14                              # rho unchanged
    np.testing.assert_allclose(out[1], 1.0, atol=1e-12)   # u = 0 against u = 1
    assert out[2] < 1e-14                              # p = rho = 1


@pytest.mark.parametrize("qid", ISO4)
def test_pressure_closed_form(oracle_lib, qid):
    q = ln.by_id(qid)
    nx, ny, u_max, tau, eps, rho0 = 4, 6, 0.03, 0.65, 1e-3, 1.02
    cs = 1.0 / q.s
    f0 = ln.equilibrium_field(q, rho0 * np.ones((nx, ny)), np.zeros((nx, ny)),
                              np.zeros((nx, ny)))
    f = np.asfortranarray(f0 + eps * q.w * (q.c[0] ** 2 - cs))
    p = rho0 + eps * cs ** 2 / (1.0 + 1.0 / (2.0 * tau))
    one = np.ones((nx, ny))
    E = _planes(rho0 * one, one, one, p * one, one, one, one, one)
    out = oracle_lib.track_errors(qid, f, E, IDENT, tau, u_max)
    assert out[2] < 1e-14
    assert out[0] < 1e-14
    # the same state against the a_bar_2-only pressure (rho + eps cs^2) misses by
    # the (1 + 1/(2 tau)) factor
    E2 = _planes(rho0 * one, one, one, (rho0 + eps * cs ** 2) * one, one, one, one, one)
    assert oracle_lib.track_errors(qid, f, E2, IDENT, tau, u_max)[2] > 1e-5


@pytest.mark.parametrize("kappa", [0.1, -0.03])
def test_velocity_error_is_relative_factor(oracle_lib, kappa):
    qid = 9
    q = ln.by_id(qid)
    nx, ny, u_max = 12, 10, 0.01
    rho, ux, uy = _smooth_fields(nx, ny, 3)
    f = np.asfortranarray(ln.equilibrium_field(q, rho, (1 + kappa) * u_max * ux,
                                               (1 + kappa) * u_max * uy))
    one = np.ones((nx, ny))
    E = _planes(rho, ux, uy, rho, one, one, one, one)
    out = oracle_lib.track_errors(qid, f, E, IDENT, 0.7, u_max)
    np.testing.assert_allclose(out[1], abs(kappa), rtol=1e-10)


@pytest.mark.parametrize("qid", [9, 37])
def test_conserved_sums_closed_form(oracle_lib, qid):
    q = ln.by_id(qid)
    nx, ny, u_max, delta = 7, 5, 0.04, 0.37
    rho0, U, V = 1.03, 0.011, -0.004
    f = np.asfortranarray(ln.equilibrium_field(q, rho0 * np.ones((nx, ny)), U * np.ones((nx, ny)),
                                               V * np.ones((nx, ny))))
    E = _planes(*np.ones((8, nx, ny)))
    out = oracle_lib.track_errors(qid, f, E, IDENT, 0.7, u_max, delta)
    N = nx * ny
    np.testing.assert_allclose(out[7], delta * N * rho0, rtol=1e-14)
    np.testing.assert_allclose(out[8], delta * N * rho0 * (U + V) / u_max, rtol=1e-12)
    np.testing.assert_allclose(out[9], delta * N * rho0 * (U * U + V * V) / u_max ** 2,
                               rtol=1e-12)


def test_affine_convention(oracle_lib):
    qid = 17
    f = random_populations(qid, 9, 8, seed=4)
    rng = np.random.default_rng(8)
    P = np.asfortranarray(rng.standard_normal((9, 8, 8)))
    aff = rng.standard_normal(16)
    direct = np.asfortranarray(aff[0::2] + aff[1::2] * P)
    a = oracle_lib.track_errors(qid, f, P, aff, 0.9, 0.03, 0.5)
    b = oracle_lib.track_errors(qid, f, direct, IDENT, 0.9, 0.03, 0.5)
    np.testing.assert_allclose(a, b, rtol=1e-14)


@pytest.mark.parametrize("qid", QIDS)
def test_c_matches_numpy(oracle_lib, qid):
    q = ln.by_id(qid)
    f = random_populations(qid, 11, 6, seed=qid)
    rng = np.random.default_rng(qid)
    E = np.asfortranarray(rng.standard_normal((11, 6, 8)))
    aff = rng.standard_normal(16)
    a = oracle_lib.track_errors(qid, f, E, aff, 0.6, 0.05, 0.25)
    b = ln.track_errors(q, f, E, aff, 0.6, 0.05, 0.25)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-13)


def test_nonfinite_nodes_propagate(oracle_lib):
    f = random_populations(9, 6, 5, seed=2)
    f[2, 3, 4] = np.nan
    E = np.asfortranarray(np.ones((6, 5, 8)))
    out = oracle_lib.track_errors(9, f, E, IDENT, 0.7, 0.02)
    assert out[10] == 1
    assert np.all(np.isnan(out[:10]))
    s = oracle_lib.velocity_stats(9, f)
    assert s[4] == 1 and np.isnan(s[0])


@pytest.mark.parametrize("qid", [5, 9, 21])
def test_velocity_stats_closed_form(oracle_lib, qid):
    q = ln.by_id(qid)
    nx, ny, U, V = 6, 7, 0.012, -0.003
    f = np.asfortranarray(ln.equilibrium_field(q, np.ones((nx, ny)), U * np.ones((nx, ny)),
                                               V * np.ones((nx, ny))))
    N = nx * ny
    u_old = np.zeros((nx, ny, 2), order="F")
    s1 = oracle_lib.velocity_stats(qid, f, u_old)
    np.testing.assert_allclose(s1[:2], [U, V], rtol=1e-13)
    np.testing.assert_allclose(s1[2], N * (U * U + V * V), rtol=1e-12)
    assert s1[3] == 0.0
    np.testing.assert_allclose(u_old[..., 0], U, rtol=1e-13)
    s2 = oracle_lib.velocity_stats(qid, f, u_old)
    assert s2[2] < 1e-28
    np.testing.assert_allclose(s2[3], N * (U * U + V * V), rtol=1e-12)
    f3 = np.asfortranarray(ln.equilibrium_field(q, np.ones((nx, ny)), 3 * U * np.ones((nx, ny)),
                                                3 * V * np.ones((nx, ny))))
    s3 = oracle_lib.velocity_stats(qid, f3, u_old)
    np.testing.assert_allclose(s3[2], 4 * N * (U * U + V * V), rtol=1e-10)
    ref = ln.velocity_stats(q, f3, np.full((nx, ny, 2), np.nan))
    np.testing.assert_allclose(s3[0:2], ref[0:2], rtol=1e-13)
