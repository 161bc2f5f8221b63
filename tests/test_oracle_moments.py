"""Pins the oracle's moment fields rho, u, sigma (moments.jl:1-19, 81-96).

* sigma of an equilibrium state is zero: the truncated equilibrium reproduces
  a_2 = rho u u exactly on quadratures with a second-order equilibrium;
* a shear perturbation f_i += eps w_i c_ix c_iy changes neither rho nor u and
  gives sigma_xy = eps c_s^4 / (1 + 1/(2 tau)) (fourth-moment isotropy
  sum_i w_i c_ix^2 c_iy^2 = c_s^4), with c_s^2 = 1/s in the reference's naming;
* the C restatement against the independent NumPy formulation.
"""
import numpy as np
import pytest

from lbm_cases import QIDS, random_populations
from oracle import lbm_numpy as ln


@pytest.mark.parametrize("qid", [9, 13, 17, 21, 37])
def test_sigma_of_equilibrium_is_zero(oracle_lib, qid):
    q = ln.by_id(qid)
    rng = np.random.default_rng(qid)
    rho = 1.0 + 0.01 * rng.standard_normal((5, 4))
    ux, uy = 0.03 * rng.standard_normal((2, 5, 4))
    f = np.asfortranarray(ln.equilibrium_field(q, rho, ux, uy))
    r, u, v, sxx, sxy, syy = oracle_lib.moments(qid, f, 0.8)
    np.testing.assert_allclose(r, rho, rtol=1e-14)
    np.testing.assert_allclose(u, ux, rtol=0, atol=1e-15)
    np.testing.assert_allclose(v, uy, rtol=0, atol=1e-15)
    for s in (sxx, sxy, syy):
        assert np.abs(s).max() < 1e-15


@pytest.mark.parametrize("qid", [9, 13, 17, 37])
def test_sigma_shear_closed_form(oracle_lib, qid):
    q = ln.by_id(qid)
    tau, eps = 0.7, 1e-4
    f = np.asfortranarray(ln.equilibrium_field(q, np.ones((3, 3)), np.zeros((3, 3)),
                                               np.zeros((3, 3))))
    f = np.asfortranarray(f + eps * q.w * q.c[0] * q.c[1])
    r, u, v, sxx, sxy, syy = oracle_lib.moments(qid, f, tau)
    np.testing.assert_allclose(r, 1.0, rtol=1e-15)
    assert np.abs(u).max() < 1e-16 and np.abs(v).max() < 1e-16
    expect = eps / q.s ** 2 / (1.0 + 1.0 / (2.0 * tau))
    np.testing.assert_allclose(sxy, expect, rtol=1e-10)
    assert np.abs(sxx).max() < 1e-16 and np.abs(syy).max() < 1e-16


@pytest.mark.parametrize("qid", QIDS)
def test_c_moments_match_numpy(oracle_lib, qid):
    q = ln.by_id(qid)
    f = random_populations(qid, 6, 5, seed=21)
    r, u, v, sxx, sxy, syy = oracle_lib.moments(qid, f, 0.9)
    rr, uu, vv = ln.moments_field(q, f)
    np.testing.assert_allclose(r, rr, rtol=1e-15)
    np.testing.assert_allclose(u, uu, rtol=0, atol=1e-16)
    np.testing.assert_allclose(v, vv, rtol=0, atol=1e-16)
    nxx, nxy, nyy = ln.deviatoric_field(q, f, 0.9)
    for a, b in ((sxx, nxx), (sxy, nxy), (syy, nyy)):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-15)
