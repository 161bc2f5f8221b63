"""Pins the oracle's collision operators.

* SRT at tau = 1 returns the equilibrium (test/quadrature.jl:64-96);
* TRT(tau, tau) == SRT(tau), on the uniform field and over 11 Poiseuille
  steps for tau in 0.51..1.1 (test/collision_models.jl:3-55);
* MRT(q, tau) == SRT(tau) on the uniform field, and over the Poiseuille steps
  at tau = 1 (test/collision_models.jl:57-114);
* conservation: mass always, momentum without force, momentum gain rho F
  per collision with the velocity-shift force of SRT and TRT;
* the C restatement against the independent NumPy formulation.
"""
import numpy as np
import pytest

from lbm_cases import QIDS, bc_dicts, random_populations
from oracle import lbm_numpy as ln

RTOL_JULIA = np.sqrt(np.finfo(float).eps)  # Julia's default isapprox


def _zeros_like(f):
    return np.zeros_like(f, order="F")


@pytest.mark.parametrize("qid", QIDS)
def test_srt_tau_one_is_equilibrium(oracle_lib, qid):
    f = random_populations(qid, 6, 5, seed=1)
    out = _zeros_like(f)
    oracle_lib.collide_srt(qid, f, out, 1.0)
    for x in range(6):
        for y in range(5):
            rho = oracle_lib.density(qid, f[x, y])
            u = oracle_lib.velocity(qid, f[x, y], rho)
            np.testing.assert_allclose(out[x, y], oracle_lib.equilibrium(qid, rho, u),
                                       rtol=0, atol=1e-16)


@pytest.mark.parametrize("qid", QIDS)
def test_trt_equal_times_is_srt(oracle_lib, qid):
    for f in (np.asfortranarray(np.broadcast_to(oracle_lib.quad_arrays(qid)[1],
                                                (10, 10, ln.by_id(qid).Q)).copy()),
              random_populations(qid, 7, 6, seed=2)):
        a, b = _zeros_like(f), _zeros_like(f)
        oracle_lib.collide_srt(qid, f, a, 0.8)
        oracle_lib.collide_trt(qid, f, b, 0.8, 0.8)
        np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-17)


def _poiseuille_init(qid, NX=3, NY=5, scale=1):
    """AnalyticalEquilibrium of PoiseuilleFlow(nu, scale) (poiseuille.jl:12-26,
    48-58; problems.jl:18-26): rho = 1, u_x = u_max y (1 - y) G / 2."""
    q = ln.by_id(qid)
    u_max = 0.1 / scale
    y = (np.arange(NY) + 0.5) / NY
    ux = np.tile(u_max * y * (1.0 - y) * 0.5, (NX, 1))
    return np.asfortranarray(ln.equilibrium_field(q, np.ones((NX, NY)), ux, np.zeros((NX, NY))))


@pytest.mark.parametrize("tau", [0.51, 0.6, 0.7, 0.8, 0.9, 1.0, 1.1])
def test_trt_equals_srt_poiseuille(oracle_lib, tau):
    f0 = _poiseuille_init(9)
    bcs = bc_dicts("poiseuille", 3, 5)
    Sa, Ca = f0.copy(order="F"), f0.copy(order="F")
    Sb, Cb = f0.copy(order="F"), f0.copy(order="F")
    oracle_lib.simulate(9, Sa, Ca, 0, [tau], 11, bcs=bcs)
    oracle_lib.simulate(9, Sb, Cb, 1, [tau, tau], 11, bcs=bcs)
    np.testing.assert_allclose(Sa, Sb, rtol=RTOL_JULIA)


def test_mrt_equals_srt_uniform(oracle_lib):
    _, w, *_ = oracle_lib.quad_arrays(9)
    f = np.asfortranarray(np.broadcast_to(w, (4, 4, 9)).copy())
    a, b = _zeros_like(f), _zeros_like(f)
    oracle_lib.collide_srt(9, f, a, 0.8)
    oracle_lib.collide_mrt(9, f, b, [0.8, 0.8])
    np.testing.assert_allclose(a, b, rtol=RTOL_JULIA)


def test_mrt_equals_srt_poiseuille_tau_one(oracle_lib):
    f0 = _poiseuille_init(9)
    bcs = bc_dicts("poiseuille", 3, 5)
    Sa, Ca = f0.copy(order="F"), f0.copy(order="F")
    Sb, Cb = f0.copy(order="F"), f0.copy(order="F")
    oracle_lib.simulate(9, Sa, Ca, 0, [1.0], 11, bcs=bcs)
    oracle_lib.simulate(9, Sb, Cb, 2, [1.0, 1.0], 11, bcs=bcs)
    np.testing.assert_allclose(Sa, Sb, rtol=RTOL_JULIA)


def test_mrt_differs_from_srt_off_unity(oracle_lib):
    """The reference marks MRT != SRT at tau != 1 as @test_broken: the
    regularised projection discards the non-hydrodynamic part of f."""
    f = random_populations(9, 5, 4, seed=3)
    a, b = _zeros_like(f), _zeros_like(f)
    oracle_lib.collide_srt(9, f, a, 0.8)
    oracle_lib.collide_mrt(9, f, b, [0.8, 0.8])
    assert np.abs(a - b).max() > 1e-6


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("coll", [0, 1, 2])
def test_conservation(oracle_lib, qid, coll):
    q = ln.by_id(qid)
    f = random_populations(qid, 5, 4, seed=4)
    out = _zeros_like(f)
    taus = {0: [0.7], 1: [0.7, 1.4], 2: [0.7, 0.9, 1.2, 0.8]}[coll]
    {0: lambda: oracle_lib.collide_srt(qid, f, out, taus[0]),
     1: lambda: oracle_lib.collide_trt(qid, f, out, *taus),
     2: lambda: oracle_lib.collide_mrt(qid, f, out, taus)}[coll]()
    np.testing.assert_allclose(out.sum(axis=2), f.sum(axis=2), rtol=1e-14)
    for d in range(2):
        np.testing.assert_allclose((out * q.c[d]).sum(axis=2), (f * q.c[d]).sum(axis=2),
                                   rtol=0, atol=1e-15)


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("coll", [0, 1])
def test_force_adds_rho_F(oracle_lib, qid, coll):
    """u' = u + tau F in the equilibrium gives sum_i c_i f_out = rho u + rho F
    (SRT: (1 - 1/tau) rho u + (1/tau) rho (u + tau F); TRT the same through
    its antisymmetric part with tau_a)."""
    q = ln.by_id(qid)
    f = random_populations(qid, 4, 3, seed=5)
    F = np.array([3e-5, -1e-5])
    out = _zeros_like(f)
    if coll == 0:
        oracle_lib.collide_srt(qid, f, out, 0.7, F)
    else:
        oracle_lib.collide_trt(qid, f, out, 0.7, 1.9, F)
    rho = f.sum(axis=2)
    for d in range(2):
        gain = (out * q.c[d]).sum(axis=2) - (f * q.c[d]).sum(axis=2)
        np.testing.assert_allclose(gain, rho * F[d], rtol=1e-9, atol=1e-16)


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("coll", [0, 1, 2])
@pytest.mark.parametrize("force", [None, "const", "field"])
def test_c_matches_numpy(oracle_lib, qid, coll, force):
    q = ln.by_id(qid)
    nx, ny = 6, 5
    f = random_populations(qid, nx, ny, seed=6)
    F = None if force is None else (np.array([1e-5, 2e-6]) if force == "const" else
                                   np.asfortranarray(1e-5 * np.random.default_rng(1)
                                                     .standard_normal((nx, ny, 2))))
    out = _zeros_like(f)
    if coll == 0:
        oracle_lib.collide_srt(qid, f, out, 0.75, F)
        ref = ln.collide_srt(q, f, 0.75, F)
    elif coll == 1:
        oracle_lib.collide_trt(qid, f, out, 0.75, 1.2, F)
        ref = ln.collide_trt(q, f, 0.75, 1.2, F)
    else:
        taus = [0.75, 0.9, 1.1, 0.6]
        oracle_lib.collide_mrt(qid, f, out, taus, F)
        ref = ln.collide_mrt(q, f, taus, F)
    np.testing.assert_allclose(out, ref, rtol=0, atol=2e-15)
