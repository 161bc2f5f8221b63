"""Pins the oracle's full step loop against analytic solutions.

* C1 shear wave: D2Q9 SRT, 64 x 64, tau = 0.8, 1000 steps; the mode decays as
  exp(-nu k^2 t) (decaying_shear_flow.jl:76-93), mass is conserved;
* Couette flow is exact on D2Q9 (README.md: "Couette flow is exact"; BCs of
  couette_flow.jl:82-85): linear profile to rounding;
* Poiseuille flow (poiseuille.jl:48-58, 72-96): parabolic steady profile;
* Taylor-Green vortex decay exp(-2 nu k^2 t) with TRT (second_order_convergence.jl:68-84);
* second-order convergence of the shear wave under diffusive scaling (the
  O(N^-2) reference lines of the shear_wave / taylor_green_vortex notebooks).
"""
import numpy as np
import pytest

from lbm_cases import shear_wave_fields, taylor_green_fields
from oracle import lbm_numpy as ln


@pytest.fixture(autouse=True)
def _threads(oracle_lib):
    oracle_lib.set_threads(4)
    yield
    oracle_lib.set_threads(1)


def _eq(qid, rho, ux, uy):
    return np.asfortranarray(ln.equilibrium_field(ln.by_id(qid), rho, ux, uy))


def _run(oracle_lib, qid, f0, coll, taus, steps, force=None, bcs=()):
    S, C = f0.copy(order="F"), f0.copy(order="F")
    oracle_lib.simulate(qid, S, C, coll, taus, steps, force=force, bcs=list(bcs))
    return S


def test_shear_wave_c1(oracle_lib):
    n, tau, steps = 64, 0.8, 1000
    f0 = _eq(9, *shear_wave_fields(n, 0.01))
    S = _run(oracle_lib, 9, f0, 0, [tau], steps)
    rho, ux, uy = oracle_lib.moments(9, S, tau)[:3]
    y = np.arange(n)
    amp = 2.0 / n * np.sum(ux[0] * np.sin(2 * np.pi * y / n))
    nu, k = (tau - 0.5) / 3.0, 2 * np.pi / n
    expect = 0.01 * np.exp(-nu * k * k * steps)
    assert abs(expect - 0.0038143) < 1e-6        # the analytic amplitude ratio 0.38143
    assert abs(amp / expect - 1.0) < 2e-3, (amp, expect)
    assert abs(rho.sum() / f0.sum() - 1.0) < 1e-13
    assert np.abs(uy).max() < 1e-13


def test_couette_exact_d2q9(oracle_lib):
    nx, ny, uw = 1, 8, 0.01
    f0 = _eq(9, np.ones((nx, ny)), np.zeros((nx, ny)), np.zeros((nx, ny)))
    bcs = [dict(kind=0, direction=2, xs=(1, nx), ys=(1, ny)),
           dict(kind=1, direction=0, xs=(1, nx), ys=(1, ny), u=(uw, 0.0), rho=1.0)]
    S = _run(oracle_lib, 9, f0, 0, [1.0], 8000, bcs=bcs)
    ux = oracle_lib.moments(9, S, 1.0)[1][0]
    exact = uw * (np.arange(ny) + 0.5) / ny      # walls half a node outside
    np.testing.assert_allclose(ux, exact, rtol=0, atol=1e-12 * uw)


def test_poiseuille_parabola(oracle_lib):
    nx, ny, F, tau = 3, 16, 1e-6, 1.0
    nu = (tau - 0.5) / 3.0
    f0 = _eq(9, np.ones((nx, ny)), np.zeros((nx, ny)), np.zeros((nx, ny)))
    bcs = [dict(kind=0, direction=0, xs=(1, nx), ys=(1, ny)),
           dict(kind=0, direction=2, xs=(1, nx), ys=(1, ny))]
    S = _run(oracle_lib, 9, f0, 0, [tau], 40000, force=np.array([F, 0.0]), bcs=bcs)
    ux = oracle_lib.moments(9, S, tau)[1][1]
    y = np.arange(ny) + 0.5
    coef = np.polyfit(y, ux, 2)
    assert np.abs(np.polyval(coef, y) - ux).max() < 1e-10 * ux.max()  # exactly quadratic
    assert abs(coef[0] / (-F / (2 * nu)) - 1.0) < 1e-6                  # curvature -F/(2 nu)
    assert np.allclose(ux, ux[::-1], rtol=1e-12)                      # symmetric about the centre


def test_taylor_green_decay_trt(oracle_lib):
    n, tau, steps = 64, 0.6, 500
    f0 = _eq(9, *taylor_green_fields(n, 0.02))
    tau_a = 0.5 + 0.25 / (tau - 0.5)
    S = _run(oracle_lib, 9, f0, 1, [tau, tau_a], steps)
    ux = oracle_lib.moments(9, S, tau)[1]
    c = (np.arange(n) + 0.5) * (2 * np.pi / n)
    X, Y = np.meshgrid(c, c, indexing="ij")
    mode = np.sin(X) * np.cos(Y)
    amp = np.sum(ux * mode) / np.sum(mode * mode)
    nu, k = (tau - 0.5) / 3.0, 2 * np.pi / n
    expect = 0.02 * np.exp(-2 * nu * k * k * steps)
    assert abs(amp / expect - 1.0) < 3e-3, (amp, expect)


def test_shear_wave_second_order_convergence(oracle_lib):
    """Diffusive scaling: N = 16, 32, 64; u0 = 0.02 * 16 / N; steps ~ N^2;
    relative L2 error against A exp(-nu k^2 t) sin(k y) falls as N^-2."""
    tau = 0.8
    nu = (tau - 0.5) / 3.0
    errs = []
    for n in (16, 32, 64):
        u0 = 0.02 * 16 / n
        steps = 150 * (n // 16) ** 2
        rho, ux, uy = shear_wave_fields(n, u0)
        S = _run(oracle_lib, 9, _eq(9, rho, ux, uy), 0, [tau], steps)
        got = oracle_lib.moments(9, S, tau)[1]
        k = 2 * np.pi / n
        exact = np.broadcast_to(u0 * np.exp(-nu * k * k * steps) * np.sin(k * np.arange(n)),
                                got.shape)
        errs.append(np.sqrt(np.sum((got - exact) ** 2) / np.sum(exact ** 2)))
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all((orders > 1.8) & (orders < 2.3)), (errs, orders)
