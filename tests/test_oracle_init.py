"""Pins of the oracle's analytic initial conditions (lbo_init_node /
lbo_initialize), initial_conditions/analytical_equilibrium.jl and
analytical_offequilibrium.jl.

* mode 0 (AnalyticalEquilibrium = the Hermite sum of hermite.jl:10-33): at
  T = 1 it is the closed-form equilibrium; at T != 1 its moments are the
  Maxwellian's to the quadrature's order: sum f = rho, sum c f = rho u,
  sum c c f = rho (u u + cs T I) (cs = 1/s), and for the order-3 sets also
  the third moment rho (u u u + cs T sym(u I));
* modes 1 and 2 (equilibrium + off-equilibrium): the added term carries no
  mass or momentum (sum_i w_i H2_i = 0, sum_i w_i c_i H2_i = 0) and its second
  Hermite moment is fixed by the orthogonality of the quadrature,
  sum_i w_i H2_i (H2_i : S) = 2 cs^2 S for symmetric S, so
  Pi^neq = sum f H2 - a_eq2 = -cs tau_eff rho S (mode 1) or
  -factor (tau + 1/2) cs S (mode 2), S = G + G^T.  With tau = s nu the
  deviatoric stress deviatoric_tensor(q, s nu, f, rho, u) (moments.jl:81-96)
  of a mode-1 start is the Newtonian -rho nu S exactly.
A dropped factor, a wrong sign, G in place of G + G^T for a non-symmetric G,
or a missing temperature term fails one of these.
"""
import numpy as np
import pytest

from lbm_cases import QIDS
from oracle import lbm_numpy as ln

ORDER2 = [9, 13, 17, 21, 37]


def _hermite2(q):
    cs = 1.0 / q.s
    c = q.c.astype(float)
    return np.stack([np.outer(c[:, i], c[:, i]) - cs * np.eye(2) for i in range(q.Q)])


@pytest.mark.parametrize("qid", QIDS)
def test_unit_temperature_is_the_closed_form(oracle_lib, qid):
    q = ln.by_id(qid)
    rng = np.random.default_rng(qid)
    for _ in range(4):
        rho = 1.0 + 0.05 * rng.standard_normal()
        u = 0.05 * rng.standard_normal(2)
        f = oracle_lib.init_node(qid, 0, rho, u)
        np.testing.assert_allclose(f, oracle_lib.equilibrium(qid, rho, u), rtol=0, atol=2e-16)


@pytest.mark.parametrize("qid", ORDER2)
@pytest.mark.parametrize("T", [0.97, 1.0, 1.04])
def test_temperature_moments(oracle_lib, qid, T):
    q = ln.by_id(qid)
    cs = 1.0 / q.s
    c = q.c.astype(float)
    rho, u = 1.02, np.array([0.03, -0.02])
    f = oracle_lib.init_node(qid, 0, rho, u, T=T)
    np.testing.assert_allclose(f.sum(), rho, rtol=1e-14)
    np.testing.assert_allclose(c @ f, rho * u, rtol=0, atol=1e-15)
    M2 = np.einsum("i,ai,bi->ab", f, c, c)
    np.testing.assert_allclose(M2, rho * (np.outer(u, u) + cs * T * np.eye(2)), rtol=0, atol=2e-15)
    if q.order >= 7:
        M3 = np.einsum("i,ai,bi,gi->abg", f, c, c, c)
        d = np.eye(2)
        sym = (np.einsum("a,bg->abg", u, d) + np.einsum("b,ag->abg", u, d)
               + np.einsum("g,ab->abg", u, d))
        want = rho * (np.einsum("a,b,g->abg", u, u, u) + cs * T * sym)
        np.testing.assert_allclose(M3, want, rtol=0, atol=2e-15)


def _offeq_cases(qid):
    rng = np.random.default_rng(100 + qid)
    for _ in range(3):
        rho = 1.0 + 0.03 * rng.standard_normal()
        u = 0.04 * rng.standard_normal(2)
        G = 1e-3 * rng.standard_normal((2, 2))          # not symmetric
        yield rho, u, G


@pytest.mark.parametrize("qid", ORDER2)
@pytest.mark.parametrize("mode", [1, 2])
def test_off_equilibrium_stress(oracle_lib, qid, mode):
    q = ln.by_id(qid)
    cs = 1.0 / q.s
    H2 = _hermite2(q)
    c = q.c.astype(float)
    tau, factor = 0.3, 1.7
    for rho, u, G in _offeq_cases(qid):
        f0 = oracle_lib.init_node(qid, 0, rho, u)
        f = oracle_lib.init_node(qid, mode, rho, u, G=G, tau=tau, factor=factor)
        d = f - f0
        np.testing.assert_allclose(d.sum(), 0.0, rtol=0, atol=1e-15)
        np.testing.assert_allclose(c @ d, 0.0, rtol=0, atol=1e-15)
        S = G + G.T
        Pi = np.einsum("i,iab->ab", d, H2)
        want = -cs * (tau + 0.5) * rho * S if mode == 1 else -factor * (tau + 0.5) * cs * S
        np.testing.assert_allclose(Pi, want, rtol=1e-11, atol=1e-15 * np.abs(want).max())


@pytest.mark.parametrize("qid", ORDER2)
def test_off_equilibrium_start_has_newtonian_stress(oracle_lib, qid):
    """deviatoric_tensor(q, s nu, f, rho, u) of the mode-1 start = -rho nu S
    (the reference's tau = s nu; tau_eff = tau + 1/2)."""
    q = ln.by_id(qid)
    nu = 0.05
    tau = q.s * nu
    nx, ny = 3, 2
    rng = np.random.default_rng(7)
    rho = 1.0 + 0.02 * rng.standard_normal((nx, ny))
    ux, uy = 0.03 * rng.standard_normal((2, nx, ny))
    G = 1e-3 * rng.standard_normal((nx, ny, 2, 2))
    f = oracle_lib.initialize(qid, 1, rho, ux, uy, G=G, tau=tau)
    r, vx, vy, sxx, sxy, syy = oracle_lib.moments(qid, f, tau)
    np.testing.assert_allclose(r, rho, rtol=1e-14)
    S = G + np.swapaxes(G, -1, -2)
    tr = 0.5 * (S[..., 0, 0] + S[..., 1, 1])
    np.testing.assert_allclose(sxx, -rho * nu * (S[..., 0, 0] - tr), rtol=1e-9, atol=1e-18)
    np.testing.assert_allclose(sxy, -rho * nu * S[..., 0, 1], rtol=1e-9, atol=1e-18)
    np.testing.assert_allclose(syy, -rho * nu * (S[..., 1, 1] - tr), rtol=1e-9, atol=1e-18)


@pytest.mark.parametrize("qid", [9, 37])
def test_lattice_matches_nodes(oracle_lib, qid):
    nx, ny = 4, 3
    rng = np.random.default_rng(3)
    rho = 1.0 + 0.02 * rng.standard_normal((nx, ny))
    ux, uy = 0.03 * rng.standard_normal((2, nx, ny))
    T = 1.0 + 0.01 * rng.standard_normal((nx, ny))
    G = 1e-3 * rng.standard_normal((nx, ny, 2, 2))
    f = oracle_lib.initialize(qid, 2, rho, ux, uy, T=T, G=G, tau=0.4, factor=2.0)
    for x in range(nx):
        for y in range(ny):
            g = oracle_lib.init_node(qid, 2, rho[x, y], [ux[x, y], uy[x, y]], T=T[x, y],
                                     G=G[x, y], tau=0.4, factor=2.0)
            assert np.array_equal(f[x, y], g)
