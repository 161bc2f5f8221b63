"""Pins the oracle's equilibrium (the closed forms used inside collide!).

* mass and momentum of the equilibrium for u over 11 decades
  (test/runtests.jl:56-118: density rtol 1e-10 / atol 1e-8, momentum
  atol 1e-11 / rtol 1e-6);
* the Hermite moments the truncated expansion must reproduce exactly:
  second moment for every order-2+ set, third for D2Q17/21/37, fourth for
  D2Q37 -- a dropped or mis-signed term of a2 / a3 / a4 fails these;
* agreement with the generic Hermite-sum equilibrium
  (velocity_distribution_function/hermite.jl:10-33), in C and in NumPy.
"""
import itertools

import numpy as np
import pytest

from lbm_cases import QIDS
from oracle import lbm_numpy as ln


def logspace(a, b, n):
    return 10.0 ** np.linspace(a, b, n)


def _ok_density(rho):
    return abs(rho - 1.0) <= 1e-8 + 1e-10


@pytest.mark.parametrize("qid", QIDS)
def test_mass_and_momentum(oracle_lib, qid):
    """runtests.jl:68-117, including its fallback: where the closed form's
    density misses (cancellation of the 4th-order terms of D2Q37 at |u| = 100)
    the reference's test replaces f by the Hermite-sum equilibrium (:77-86)."""
    for u in logspace(2, -9, 12):
        for v in ([0.0, 0.0], [u, 0.0], [0.0, u], [u, u]):
            f = oracle_lib.equilibrium(qid, 1.0, np.array(v))
            if u <= 1.0:
                assert _ok_density(oracle_lib.density(qid, f)), (qid, v)
            if not _ok_density(oracle_lib.density(qid, f)):
                f = oracle_lib.hermite_equilibrium(qid, 1.0, np.array(v))
            rho = oracle_lib.density(qid, f)
            assert _ok_density(rho), (qid, v, rho)
            m = oracle_lib.velocity(qid, f, 1.0)
            for d in range(2):
                assert abs(m[d] - v[d]) <= 1e-11 + 1e-6 * abs(v[d]), (qid, v, m)


def _moment(q, f, n):
    c = q.c.astype(float)
    idx = "abcd"[:n]
    ops = [f] + [c[k] for k in range(2)]
    out = np.zeros((2,) * n)
    for comp in itertools.product(range(2), repeat=n):
        out[comp] = np.sum(f * np.prod([c[a] for a in comp], axis=0))
    return out


def _expected(n, rho, u, cs):
    d = np.eye(2)
    if n == 2:
        return rho * (np.einsum("a,b->ab", u, u) + cs * d)
    if n == 3:
        uuu = np.einsum("a,b,c->abc", u, u, u)
        sym = (np.einsum("a,bc->abc", u, d) + np.einsum("b,ac->abc", u, d)
               + np.einsum("c,ab->abc", u, d))
        return rho * (uuu + cs * sym)
    uuuu = np.einsum("a,b,c,d->abcd", u, u, u, u)
    uu = np.einsum("a,b->ab", u, u)
    s2 = (np.einsum("ab,cd->abcd", uu, d) + np.einsum("ac,bd->abcd", uu, d)
          + np.einsum("ad,bc->abcd", uu, d) + np.einsum("bc,ad->abcd", uu, d)
          + np.einsum("bd,ac->abcd", uu, d) + np.einsum("cd,ab->abcd", uu, d))
    s0 = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d)
          + np.einsum("ad,bc->abcd", d, d))
    return rho * (uuuu + cs * s2 + cs * cs * s0)


@pytest.mark.parametrize("qid,nmax", [(9, 2), (13, 2), (17, 3), (21, 3), (37, 4)])
def test_hermite_moments(oracle_lib, qid, nmax):
    q = ln.by_id(qid)
    rho = 1.03
    for u in (np.array([0.05, -0.02]), np.array([0.1, 0.1]), np.array([-0.003, 0.07])):
        f = oracle_lib.equilibrium(qid, rho, u)
        for n in range(2, nmax + 1):
            got = _moment(q, f, n)
            exp = _expected(n, rho, u, 1.0 / q.s)
            np.testing.assert_allclose(got, exp, atol=2e-15, err_msg=f"D2Q{qid} moment {n}")


@pytest.mark.parametrize("qid", QIDS)
def test_closed_form_equals_hermite_sum(oracle_lib, qid):
    q = ln.by_id(qid)
    rng = np.random.default_rng(qid)
    for _ in range(5):
        rho = 1.0 + 0.05 * rng.standard_normal()
        u = 0.08 * rng.standard_normal(2)
        fe = oracle_lib.equilibrium(qid, rho, u)
        fh = oracle_lib.hermite_equilibrium(qid, rho, u)
        fn = ln.hermite_equilibrium(q, rho, u)
        np.testing.assert_allclose(fe, fh, rtol=0, atol=1e-15)
        np.testing.assert_allclose(fe, fn, rtol=0, atol=1e-15)


def test_rest_state_is_weights(oracle_lib):
    for qid in QIDS:
        _, w, *_ = oracle_lib.quad_arrays(qid)
        assert np.array_equal(oracle_lib.equilibrium(qid, 1.0, np.zeros(2)), w)
