"""Closed-form pins of the oracle's regularised MRT on every quadrature
(src/collision_models/mrt.jl:81-116).

The reference tests MRT on D2Q9 only (test/collision_models.jl:57-114).  Two
identities fix the higher-order terms on D2Q13/17/21/37 as well:

(i)  with every tau_n = 1 the collision returns the Hermite-sum equilibrium
     of order N = div(order, 2): f_i = w_i sum_n a_eq[n] : H_n(c_i) / (n! cs^n)
     with a_eq[n] = rho u^(x)n (hermite.jl:10-33, 37-77);
(ii) for distinct tau_n and every n <= N, the n-th Hermite moment of the
     collided populations is the relaxed coefficient:
     sum_i f_out,i H_n(c_i) = (1 - 1/tau_n) a_f[n] + a_eq[n] / tau_n,
     a_f[n] = sum_i f_i H_n(c_i) (mrt.jl:99-104).  This holds because the
     reconstruction w_i s^n/n! a : H_n(c_i) has n-th moment a when the
     quadrature integrates degree 2N exactly.

The Hermite tensors here come from a construction the oracle does not use:
a component with k x-indices and n - k y-indices equals
cs^(n/2) He_k(c_x / sqrt(cs)) He_{n-k}(c_y / sqrt(cs)) with He the
probabilists' Hermite polynomials (numpy.polynomial.hermite_e).  A wrong
s^n/n!, a wrong a_eq[n] or a wrong relaxation of an n >= 3 coefficient fails
(i) or (ii); the planted-mistake tests below show it for the NumPy
restatement.
"""
import math
from math import comb

import numpy as np
import pytest
from numpy.polynomial import hermite_e as He

from lbm_cases import random_populations
from oracle import lbm_numpy as ln

MRT_QIDS = [9, 13, 17, 21, 37]


def _he(k, x):
    return He.hermeval(x, [0] * k + [1])


def _hermite_components(q, n):
    """(n + 1, Q): row k is the tensor component with k x-indices."""
    cs = 1.0 / q.s
    r = np.sqrt(cs)
    cx, cy = q.c[0].astype(float) / r, q.c[1].astype(float) / r
    return np.stack([cs ** (n / 2) * _he(k, cx) * _he(n - k, cy) for k in range(n + 1)])


def _moments(q, f, n):
    """(nx, ny, n + 1) Hermite moments of order n of a population field."""
    return np.einsum("xyi,ki->xyk", f, _hermite_components(q, n))


def _rho_u(q, f):
    rho = f.sum(axis=2)
    return rho, (f * q.c[0]).sum(axis=2) / rho, (f * q.c[1]).sum(axis=2) / rho


def _aeq(rho, ux, uy, n):
    return np.stack([rho * ux ** k * uy ** (n - k) for k in range(n + 1)], axis=-1)


def hermite_sum_equilibrium(q, rho, ux, uy):
    """Identity (i): the order-N Hermite-sum equilibrium, contracted
    component-wise (a_eq is symmetric: C(n, k) equal components)."""
    cs = 1.0 / q.s
    N = q.order // 2
    out = np.zeros(rho.shape + (q.Q,))
    for n in range(N + 1):
        H = _hermite_components(q, n)
        a = _aeq(rho, ux, uy, n) * np.array([comb(n, k) for k in range(n + 1)])
        out += np.einsum("xyk,ki->xyi", a, H) / (math.factorial(n) * cs ** n)
    return out * q.w


def mrt_identity_errors(q, f, out, taus):
    """Largest violation of (ii) over n = 0..N, relative to the moment's size."""
    rho, ux, uy = _rho_u(q, f)
    errs = []
    for n in range(q.order // 2 + 1):
        got = _moments(q, out, n)
        if n <= 1:
            want = _aeq(rho, ux, uy, n)           # mass and momentum are conserved
        else:
            t = taus[n - 1]
            want = (1.0 - 1.0 / t) * _moments(q, f, n) + _aeq(rho, ux, uy, n) / t
        errs.append(np.abs(got - want).max() / max(np.abs(want).max(), 1e-300))
    return errs


def _field(qid, seed):
    # non-equilibrium parts large enough that every Hermite order is populated
    return random_populations(qid, 5, 4, seed=seed, u_amp=0.05, noise=2e-2)


@pytest.mark.parametrize("qid", MRT_QIDS)
def test_identities_hold_for_the_equilibrium(qid):
    """Sanity of the test's own Hermite construction: the closed-form
    equilibrium (pinned by its moments elsewhere) equals (i)'s sum."""
    q = ln.by_id(qid)
    rng = np.random.default_rng(qid)
    rho = 1.0 + 0.02 * rng.standard_normal((3, 2))
    ux, uy = 0.05 * rng.standard_normal((2, 3, 2))
    np.testing.assert_allclose(hermite_sum_equilibrium(q, rho, ux, uy),
                               ln.equilibrium_field(q, rho, ux, uy), rtol=0, atol=2e-16)


@pytest.mark.parametrize("qid", MRT_QIDS)
def test_mrt_unit_times_is_hermite_equilibrium(oracle_lib, qid):
    q = ln.by_id(qid)
    f = _field(qid, 11)
    taus = [1.0] * 4
    rho, ux, uy = _rho_u(q, f)
    want = hermite_sum_equilibrium(q, rho, ux, uy)
    out = np.zeros_like(f, order="F")
    oracle_lib.collide_mrt(qid, f, out, taus)
    np.testing.assert_allclose(out, want, rtol=0, atol=5e-16)
    np.testing.assert_allclose(ln.collide_mrt(q, f, taus), want, rtol=0, atol=5e-16)


@pytest.mark.parametrize("qid", MRT_QIDS)
@pytest.mark.parametrize("taus", [[0.7, 0.9, 1.3, 0.6], [1.6, 0.55, 0.8, 1.9]])
def test_mrt_relaxes_each_hermite_order(oracle_lib, qid, taus):
    q = ln.by_id(qid)
    f = _field(qid, 12)
    out = np.zeros_like(f, order="F")
    oracle_lib.collide_mrt(qid, f, out, taus)
    assert max(mrt_identity_errors(q, f, out, taus)) < 1e-12
    assert max(mrt_identity_errors(q, f, ln.collide_mrt(q, f, taus), taus)) < 1e-12


# ---- planted mistakes in the NumPy restatement: each must fail a pin ----------
class _WrongFactorial:
    """math with n! wrong for one n (the s^n/n! of the reconstruction)."""

    def __init__(self, bad_n):
        self.bad_n = bad_n

    def factorial(self, n):
        return math.factorial(n) * (0.5 if n == self.bad_n else 1.0)

    def __getattr__(self, name):
        return getattr(math, name)


@pytest.mark.parametrize("qid,bad_n", [(17, 3), (21, 3), (37, 3), (37, 4), (13, 2)])
def test_planted_wrong_factorial_fails(monkeypatch, qid, bad_n):
    q = ln.by_id(qid)
    f = _field(qid, 13)
    monkeypatch.setattr(ln, "math", _WrongFactorial(bad_n))
    taus = [0.7, 0.9, 1.3, 0.6]
    assert max(mrt_identity_errors(q, f, ln.collide_mrt(q, f, taus), taus)) > 1e-3
    rho, ux, uy = _rho_u(q, f)
    ones = [1.0] * 4
    assert np.abs(ln.collide_mrt(q, f, ones) - hermite_sum_equilibrium(q, rho, ux, uy)).max() > 1e-8


@pytest.mark.parametrize("qid", [17, 37])
def test_planted_wrong_hermite_tensor_fails(monkeypatch, qid):
    """H_3 with the cs term's sign flipped (projection and reconstruction both
    use it, so only a pin from an independent construction can see it)."""
    q = ln.by_id(qid)
    f = _field(qid, 14)
    good = ln.hermite_tensor

    def bad(n, xi, cs):
        if n != 3:
            return good(n, xi, cs)
        return good(3, xi, -cs)

    monkeypatch.setattr(ln, "hermite_tensor", bad)
    taus = [0.7, 0.9, 1.3, 0.6]
    assert max(mrt_identity_errors(q, f, ln.collide_mrt(q, f, taus), taus)) > 1e-3


@pytest.mark.parametrize("qid", [17, 37])
def test_planted_wrong_relaxation_fails(qid):
    """Relaxing order 3 with tau_2 (an index slip in a_coll) is caught by (ii)."""
    q = ln.by_id(qid)
    f = _field(qid, 15)
    taus = [0.7, 0.9, 1.3, 0.6]
    slipped = [taus[0], taus[1], taus[1], taus[3]]
    out = ln.collide_mrt(q, f, slipped)
    assert max(mrt_identity_errors(q, f, out, taus)) > 1e-3
