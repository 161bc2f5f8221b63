"""The plain FP64 Taylor-Green workload (D2Q9 TRT, magic parameter 1/4,
periodic N x N, u0 = 0.02, pressure-consistent equilibrium start, 5000
steps) through the C ABI, against the oracle and against the mathematics.

* N = 256 at tau = 0.51 (tau_a = 25.5, the stability edge), 0.6 and 1.0,
  and N = 1024 at tau = 0.6, all 5000 steps: bit-exact against the oracle
  (second_order_convergence.jl:1-131 for the vortex; trt.jl:8-21 for the
  magic-parameter factory);
* N = 4096 (the benchmark lattice), 5000 steps: size-independent properties
  -- the run split into two calls gives the same bits as one call, mass and
  momentum are conserved, the vortex's mirror symmetries hold to rounding,
  and the vortex mode follows exp(-2 nu k^2 t) within O(Ma^2);
* second-order convergence of velocity AND deviatoric stress under
  diffusive scaling (N = 64, 128, 256; u0 ~ 1/N; steps ~ N^2), starting from
  the reference's iterative initialization (mei_et_al.jl:13-41) so that the
  start carries the non-equilibrium stress; the equilibrium start adds an
  acoustic transient whose error falls only as ~1/N (measured with the
  oracle: orders 0.5-0.6 at tau = 0.6);
* Couette flow is exact on D2Q9 (README.md:13; couette_flow.jl:45-52,
  82-85) through the device path.
"""
import numpy as np
import pytest

from lbm_cases import taylor_green_fields

pytestmark = pytest.mark.gpu

ORACLE_THREADS = 16


def _L():
    import lbmhip
    return lbmhip


def _oracle_run(oracle_lib, f0, taus, steps, coll=1, bcs=()):
    S, C = f0.copy(order="F"), f0.copy(order="F")
    oracle_lib.set_threads(ORACLE_THREADS)
    try:
        oracle_lib.simulate(9, S, C, coll, list(taus), steps, bcs=list(bcs))
    finally:
        oracle_lib.set_threads(1)
    return S, C


def _tgv_f0(L, q, n, u0=0.02):
    return L.equilibrium_populations(q, *taylor_green_fields(n, u0))


def _same(a, b, what):
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        raise AssertionError(f"{what}: not bit-exact, max |diff| {d.max():.3e}, "
                             f"{np.count_nonzero(d)} of {d.size} differ")


@pytest.mark.parametrize("n,tau", [(256, 0.51), (256, 0.6), (256, 1.0), (1024, 0.6)])
def test_taylor_green_5000_steps_bit_exact(oracle_lib, n, tau):
    L = _L()
    q = L.D2Q9()
    steps = 5000
    f0 = _tgv_f0(L, q, n)
    cm = L.trt_magic(tau, 0.25)
    print(f"TGV D2Q9 TRT {n}^2 tau={tau} tau_a={cm.τ_asymmetric}: {steps} steps", flush=True)
    model = L.LatticeBoltzmannModel.from_populations(q, f0, cm)
    L.simulate(model, range(steps))
    S, C = _oracle_run(oracle_lib, f0, [cm.τ_symmetric, cm.τ_asymmetric], steps)
    _same(model.f_stream.to_numpy(), S, f"TGV {n} tau={tau} f_stream")
    _same(model.f_collision.to_numpy(), C, f"TGV {n} tau={tau} f_collision")


def test_taylor_green_4096_properties():
    L = _L()
    q = L.D2Q9()
    n, tau, steps, u0 = 4096, 0.6, 5000, 0.02
    f0 = _tgv_f0(L, q, n, u0)
    cm = L.trt_magic(tau, 0.25)
    print(f"TGV D2Q9 TRT {n}^2: {steps} steps, one call and two calls", flush=True)
    a = L.LatticeBoltzmannModel.from_populations(q, f0, cm)
    L.simulate(a, range(steps))
    b = L.LatticeBoltzmannModel.from_populations(q, f0, cm)
    L.simulate(b, range(1000))
    L.simulate(b, range(1000, steps))
    m0 = L.conserved_sums(L.HipLattice.from_numpy(q, f0)).mass
    del f0
    fa = a.f_stream.to_numpy()
    _same(fa, b.f_stream.to_numpy(), "4096 split run")
    del fa
    s = L.conserved_sums(a.f_stream)
    # FP64 rounding of the populations drifts the total by ~1e-16 per node and
    # step (the oracle-checked runs above carry the same bits); a leak would be
    # orders of magnitude larger
    assert abs(s.mass / m0 - 1.0) < 5e-12, s.mass / m0
    assert s.nonfinite_nodes == 0
    assert abs(s.momentum_x) < 1e-9 and abs(s.momentum_y) < 1e-9, s
    mom = L.moments(a.f_stream, stress=False)
    # the vortex is odd in x for u_x and even for u_y (x -> 2 pi - x maps node
    # j to n - 1 - j), and the reverse in y; the lattice step keeps both
    # mirror symmetries up to rounding
    tol = 1e-12 * u0
    assert np.abs(mom.ux + mom.ux[::-1, :]).max() < tol
    assert np.abs(mom.uy - mom.uy[::-1, :]).max() < tol
    assert np.abs(mom.ux - mom.ux[:, ::-1]).max() < tol
    assert np.abs(mom.uy + mom.uy[:, ::-1]).max() < tol
    # the vortex mode against the incompressible decay exp(-2 nu k^2 t): at
    # fixed u0 the weakly compressible lattice fluid exchanges O(Ma^2) of the
    # mode with the acoustic modes of the equilibrium start (measured with
    # the oracle: an excess damping that scales as u0^2), so the bound is a
    # three Ma^2 (Ma^2 = 3 u0^2) -- the viscous decrement over 5000 steps at this
    # size is itself only 7.8e-4
    c = (np.arange(n) + 0.5) * (2 * np.pi / n)
    sx, cx = np.sin(c), np.cos(c)
    amp = (sx @ mom.ux @ cx) / ((sx @ sx) * (cx @ cx))
    nu, k = (tau - 0.5) / 3.0, 2 * np.pi / n
    expect = u0 * np.exp(-2 * nu * k * k * steps)
    assert abs(amp / expect - 1.0) < 3 * 3 * u0 * u0, (amp, expect)


class _Vortex:
    """The lattice velocity of the Taylor-Green vortex for the iterative
    initialization (problems.jl:127-134, lattice units)."""

    def __init__(self, n, u0):
        self.NX = self.NY = n
        self.u0 = u0

    def lattice_velocity_field(self):
        _, ux, uy = taylor_green_fields(self.NX, self.u0)
        return np.asfortranarray(np.stack([ux, uy], axis=-1))


def _convergence_errors(L, tau, sizes=(64, 128, 256), u0_64=0.02, c=0.5):
    q = L.D2Q9()
    cm = L.trt_magic(tau, 0.25)
    nu = (tau - 0.5) / 3.0
    eu, es = [], []
    for n in sizes:
        u0 = u0_64 * 64 / n
        steps = int(round(c * n * n))
        f, _ = L.initialize_mei(L.IterativeInitializationMeiEtAl(1.0, 1e-12), q, _Vortex(n, u0),
                                max_steps=200000)
        model = L.LatticeBoltzmannModel.from_populations(q, f.to_numpy(), cm)
        L.simulate(model, range(steps))
        # deviatoric stress with the reference's tau = s nu (moments.jl:81-96)
        m = L.moments(model.f_stream, 3.0 * nu)
        k = 2 * np.pi / n
        cc = (np.arange(n) + 0.5) * k
        X, Y = np.meshgrid(cc, cc, indexing="ij")
        d = np.exp(-2 * nu * k * k * steps)
        ex, ey = u0 * d * np.sin(X) * np.cos(Y), -u0 * d * np.cos(X) * np.sin(Y)
        # sigma = -rho nu (grad u + grad u^T): xx = -2 nu u0 k cos x cos y = -yy, xy = 0
        sxx = -2 * nu * u0 * k * d * np.cos(X) * np.cos(Y)
        eu.append(np.sqrt(np.sum((m.ux - ex) ** 2 + (m.uy - ey) ** 2) / np.sum(ex ** 2 + ey ** 2)))
        es.append(np.sqrt(np.sum((m.sigma_xx - sxx) ** 2 + (m.sigma_yy + sxx) ** 2 + 2 * m.sigma_xy ** 2)
                          / np.sum(2 * sxx ** 2)))
        print(f"tau={tau} N={n}: {steps} steps, velocity error {eu[-1]:.3e}, "
              f"stress error {es[-1]:.3e}", flush=True)
    return np.array(eu), np.array(es)


@pytest.mark.parametrize("tau", [0.6, 1.0])
def test_second_order_convergence_velocity_and_stress(tau):
    L = _L()
    eu, es = _convergence_errors(L, tau)
    ou = np.log2(eu[:-1] / eu[1:])
    os_ = np.log2(es[:-1] / es[1:])
    assert np.all((ou > 1.8) & (ou < 2.3)), (eu, ou)
    assert np.all((os_ > 1.8) & (os_ < 2.3)), (es, os_)


def test_couette_exact_d2q9(oracle_lib):
    """Bounce-back bottom, moving top wall u_w: the steady profile is linear
    through the half-way walls, exact to rounding on D2Q9."""
    L = _L()
    q = L.D2Q9()
    nx, ny, uw, steps = 1, 8, 0.01, 8000
    f0 = L.equilibrium_populations(q, np.ones((nx, ny)), np.zeros((nx, ny)), np.zeros((nx, ny)))
    full = ((1, nx), (1, ny))
    bcs = [L.BounceBack(L.South(), *full), L.MovingWall(L.North(), *full, (uw, 0.0))]
    model = L.LatticeBoltzmannModel.from_populations(q, f0, L.SRT(1.0), bcs)
    L.simulate(model, range(steps))
    bcd = [dict(kind=0, direction=2, xs=(1, nx), ys=(1, ny)),
           dict(kind=1, direction=0, xs=(1, nx), ys=(1, ny), u=(uw, 0.0), rho=1.0)]
    S, _ = _oracle_run(oracle_lib, f0, [1.0], steps, coll=0, bcs=bcd)
    _same(model.f_stream.to_numpy(), S, "Couette")
    ux = L.moments(model.f_stream, 0.5, stress=False).ux[0]
    exact = uw * (np.arange(ny) + 0.5) / ny
    np.testing.assert_allclose(ux, exact, rtol=0, atol=1e-12 * uw)
