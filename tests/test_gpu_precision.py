"""Uniform acceleration of a fully periodic lattice in the FP32-arithmetic
storages: every node holds the same state, so streaming only permutes equal
values and the velocity after N steps is N F exactly (FP64 reproduces it to
1e-12).  What the FP32 storages lose or gain is the collision's arithmetic:
a momentum defect e of the FP32 equilibrium adds omega e rho u per step,
a relative error of about omega e N / 2.  With the D2Q17 weights and s
rounded one by one (e = -2.8e-8) FP32 ended 20,000 steps at -5.3e-4 and
FP16-NEQ at -7.3e-4; with the FP32 constants of lbm_quadrature_f32 they end
at -3.3e-4 and -2.8e-4.  The bound sits between the two.
(SRT with a constant body force: srt.jl:18-62; D2Q17: quadratures/D2Q17.jl.)
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 20000
F = 1e-6


def _run(L, q, storage, n=64):
    f0 = L.equilibrium_populations(q, np.ones((n, n)), np.zeros((n, n)), np.zeros((n, n)))
    cm = L.SRT(0.8, np.array([F, 0.0]))
    m = L.LatticeBoltzmannModel.from_populations(q, f0, cm, [], storage=storage)
    L.simulate(m, range(STEPS))
    mom = L.moments(m.f_stream, stress=False)
    return mom.ux, mom.uy


def test_uniform_acceleration_d2q17_fp32_storages():
    import lbmhip as L
    q = L.Quadratures["D2Q17"]()
    print(f"D2Q17 periodic 64x64, SRT 0.8, F={F}, {STEPS} steps", flush=True)
    ux64, uy64 = _run(L, q, "f64")
    assert abs(ux64.mean() / (STEPS * F) - 1.0) < 1e-9
    for storage in ("f32", "f16neq"):
        ux, uy = _run(L, q, storage)
        # spatially uniform to rounding, and no transverse drift
        assert np.ptp(ux) <= 1e-6 * abs(ux.mean())
        assert np.abs(uy).max() < 1e-9
        rel = ux.mean() / ux64.mean() - 1.0
        print(f"  {storage}: relative velocity error {rel:+.3e}", flush=True)
        assert abs(rel) < 4.3e-4, (storage, rel)
