"""Host logic of the on-device processing (no GPU): the Taylor-Green vortex's
analytic fields and their plane x factor split, the stopping criteria's
decisions (stopping_criteria.jl:17-115), TrackHydrodynamicErrors' control
flow (track_hydrodynamic_errors.jl:52-66) and the fused chunks simulate runs
between the steps a processing method reads."""
import math

import numpy as np
import pytest

import lbmhip as L
from lbmhip import model as model_mod
from lbmhip import processing as P


@pytest.mark.parametrize("static", [True, False])
@pytest.mark.parametrize("t", [0.0, 3.7])
def test_tgv_planes_times_factors_equal_fields(static, t):
    pb = L.TaylorGreenVortex(scale=1, NX=24, static=static)
    q = L.D2Q9()
    planes = pb.expected_planes(q)
    aff = np.asarray(pb.expected_affine(q, t)).reshape(8, 2)
    np.testing.assert_allclose(aff[:, 0] + aff[:, 1] * planes, pb.expected_fields(q, t),
                               rtol=1e-14, atol=1e-15)


def test_tgv_decay_and_stress_from_velocity():
    pb = L.TaylorGreenVortex(scale=2, static=False)
    q = L.D2Q9()
    # velocity decays as exp(-(a^2+b^2) nu t), pressure deviation as its square
    t = 5.0
    d = math.exp(-2 * pb.viscosity() * t)
    ux0, _ = pb.velocity(0.0)
    ux, _ = pb.velocity(t)
    np.testing.assert_allclose(ux, d * ux0, rtol=1e-14, atol=1e-16)
    p0, p = pb.pressure(q, 0.0) - 1, pb.pressure(q, t) - 1
    np.testing.assert_allclose(p, d * d * p0, rtol=1e-12, atol=1e-18)
    # the stress is -nu (grad u + grad u^T) of the velocity field: central
    # differences of the analytic velocity at the cell centres
    x, y = pb.range()
    h = 1e-6
    pbx = L.TaylorGreenVortex(scale=2, static=False)
    sxx, sxy, syx, syy = pb.deviatoric_tensor(q, t)
    X, Y = np.meshgrid(x, y, indexing="ij")

    def vel(Xs, Ys):
        return (d * (pbx.A * np.cos(pbx.a * Xs) * np.sin(pbx.b * Ys)),
                d * (pbx.B * np.sin(pbx.a * Xs) * np.cos(pbx.b * Ys)))
    dudx = (vel(X + h, Y)[0] - vel(X - h, Y)[0]) / (2 * h)
    dudy = (vel(X, Y + h)[0] - vel(X, Y - h)[0]) / (2 * h)
    dvdx = (vel(X + h, Y)[1] - vel(X - h, Y)[1]) / (2 * h)
    dvdy = (vel(X, Y + h)[1] - vel(X, Y - h)[1]) / (2 * h)
    nu = pb.viscosity()
    np.testing.assert_allclose(sxx, -nu * 2 * dudx, atol=1e-9)
    np.testing.assert_allclose(sxy, -nu * (dudy + dvdx), atol=1e-9)
    np.testing.assert_allclose(syx, sxy, rtol=0, atol=0)
    np.testing.assert_allclose(syy, -nu * 2 * dvdy, atol=1e-9)
    # and incompressible: the trace vanishes
    np.testing.assert_allclose(sxx + syy, 0.0, atol=1e-15)


def test_tgv_scalings():
    pb = L.TaylorGreenVortex()              # ν = 1/6, scale 2, NX = NY = 32
    assert pb.NX == 32 and pb.u_max == 0.005
    np.testing.assert_allclose(pb.delta_x(), 2 * math.pi / 32)
    np.testing.assert_allclose(pb.delta_t(), pb.delta_x() * 0.005)
    np.testing.assert_allclose(pb.viscosity(), (1 / 6) * pb.delta_x() ** 2 / pb.delta_t())
    x, y = pb.range()
    np.testing.assert_allclose(x[0], math.pi / 32)
    np.testing.assert_allclose(x[-1], 2 * math.pi - math.pi / 32)
    np.testing.assert_allclose(np.diff(x), pb.delta_x())


def test_stop_criteria_mapping():
    class Pb:
        pass
    p = Pb()
    for kind, tol in [("poiseuille", 1e-12), ("couette", 1e-7), ("lid_driven_cavity", 1e-5)]:
        p.kind = kind
        c = P.stop_criteria_for(p)
        assert isinstance(c, P.MeanVelocityStoppingCriteria) and c.tolerance == tol
    p.kind, p.static = "decaying_shear", True
    assert P.stop_criteria_for(p).tolerance == 1e-8
    p.static = False
    assert isinstance(P.stop_criteria_for(p), P.NoStoppingCriteria)
    assert isinstance(P.stop_criteria_for(L.TaylorGreenVortex()), P.NoStoppingCriteria)


def test_mean_velocity_criterion_decisions(monkeypatch):
    seq = iter([0.0100, 0.0100 * (1 + 1e-9), 0.0101, float("nan")])
    monkeypatch.setattr(P, "velocity_stats", lambda f, tracker=None: np.array([next(seq), 0, 0, 0, 0]))
    c = P.MeanVelocityStoppingCriteria(0.0, 1e-8)
    assert c.should_stop_(None, None) is False      # first check: u/0 - 1 = inf
    assert c.old_mean_velocity == 0.0100
    assert c.should_stop_(None, None) is True       # relative change 1e-9 < 1e-8
    c2 = P.MeanVelocityStoppingCriteria(0.0100, 1e-8)
    assert c2.should_stop_(None, None) is False     # 1% change
    assert c2.should_stop_(None, None) is True      # NaN


def test_velocity_convergence_criterion_decisions(monkeypatch):
    sums = iter([(4.0, 0.0), (1e-20, 2.0), (1.0, 2.0), (float("nan"), 2.0)])

    def fake(f, tracker=None):
        du2, old2 = next(sums)
        return np.array([0.0, 0.0, du2, old2, 0.0])
    monkeypatch.setattr(P, "velocity_stats", fake)
    monkeypatch.setattr(P, "VelocityTracker", lambda nx, ny: object())
    c = P.VelocityConvergenceStoppingCriteria(1e-6, nx=4, ny=4)
    assert c.should_stop_(None, None) is False      # sqrt(4) / 0 = inf
    assert c.should_stop_(None, None) is True       # 1e-10 / 2 < 1e-6
    assert c.should_stop_(None, None) is False      # 1 / 2
    assert c.should_stop_(None, None) is True       # NaN


class _Recorder(P.TrackHydrodynamicErrors):
    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self.rows_at = []

    def row(self, q, f, t):
        self.rows_at.append(t)
        return P.TrackRow(t, *([0.0] * 10))


class _StopAt(P.StopCriteria):
    def __init__(self, t_stop):
        self.t_stop, self.calls = t_stop, []

    def should_stop_(self, q, f):
        self.calls.append(len(self.calls))
        return 100 * len(self.calls) >= self.t_stop


def test_track_next_control_flow():
    pm = _Recorder(L.TaylorGreenVortex(), should_process=False, n_steps=250,
                   stop_criteria=_StopAt(10 ** 9))
    for t in range(1, 252):
        assert pm.next_(None, None, t) is False
        assert pm.needs_step(t) == (t % 100 == 0 or t == 250)
    assert pm.rows_at == [250]                      # only the last step records a row
    pm2 = _Recorder(L.TaylorGreenVortex(), should_process=False, n_steps=1000,
                    stop_criteria=_StopAt(300))
    out = [pm2.next_(None, None, t) for t in range(1, 301)]
    assert out[-1] is True and not any(out[:-1])    # stops at the third check
    assert pm2.rows_at == [300]
    pm3 = _Recorder(L.TaylorGreenVortex(), should_process=True, n_steps=5,
                    stop_criteria=_StopAt(10 ** 9))
    for t in range(1, 6):
        pm3.next_(None, None, t)
    assert pm3.rows_at == [1, 2, 3, 4, 5] and all(pm3.needs_step(t) for t in range(1, 6))


class _FakeLib:
    def __init__(self):
        self.counts = []

    def lbm_simulate(self, h, fs, fc, count):
        self.counts.append(int(count))
        return 0


class _FakeNative:
    handle = None

    def configure(self, *a, **k):
        return self

    def set_launch(self, launch):
        pass


class _FakeLattice:
    handle = None


def test_simulate_runs_fused_chunks_between_processing_steps(monkeypatch):
    pm = _Recorder(L.TaylorGreenVortex(), should_process=False, n_steps=250,
                   stop_criteria=_StopAt(10 ** 9))
    m = L.LatticeBoltzmannModel(_FakeLattice(), _FakeLattice(), L.D2Q9(), L.SRT(0.8),
                                processing_method=pm)
    fake = _FakeLib()
    monkeypatch.setattr(model_mod._lib, "lib", lambda: fake)
    monkeypatch.setattr(model_mod, "native_model", lambda q: _FakeNative())
    L.simulate(m, range(0, 251))                    # Julia 0:250
    assert fake.counts == [100, 100, 50, 1]
    assert sum(fake.counts) == 251
    assert pm.rows_at == [250]
    # a stop at the second check ends the run there
    fake.counts.clear()
    pm2 = _Recorder(L.TaylorGreenVortex(), should_process=False, n_steps=10 ** 6,
                    stop_criteria=_StopAt(200))
    m.processing_method = pm2
    L.simulate(m, range(0, 1000))
    assert fake.counts == [100, 100] and pm2.rows_at == [200]
