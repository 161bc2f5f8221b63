"""Host-side mirror of the reference API (no GPU needed)."""
import numpy as np
import pytest


def test_trt_two_argument_swap():
    """TRT(a, b) = TRT(tau_s = b, tau_a = a) (trt.jl:6)."""
    import lbmhip as L
    t = L.TRT(0.9, 1.1)
    assert t.τ_symmetric == 1.1 and t.τ_asymmetric == 0.9
    u = L.TRT.fields(0.9, 1.1)
    assert u.τ_symmetric == 0.9 and u.τ_asymmetric == 1.1


def test_trt_magic_parameter():
    """tau_a = 1/2 + Lambda / (tau - 1/2) (trt.jl:14-15)."""
    import lbmhip as L
    t = L.trt_magic(0.6, 0.25)
    assert t.τ_symmetric == 0.6 and t.τ_asymmetric == pytest.approx(3.0)
    assert L.trt_magic(1.0).τ_asymmetric == pytest.approx(1.0)
    assert L.trt_magic(0.51).τ_asymmetric == pytest.approx(25.5)


def test_collision_model_from_viscosity():
    """tau = s nu + 1/2 (srt.jl:7-16, trt.jl:8-21, mrt.jl:36-47)."""
    import lbmhip as L
    q = L.D2Q9()
    assert L.collision_model_from_viscosity(L.SRT, q, 0.1).τ == pytest.approx(0.8)
    t = L.collision_model_from_viscosity(L.TRT, q, 0.1)
    assert t.τ_symmetric == pytest.approx(0.8)
    assert L.collision_model_from_viscosity(L.TRT_Λ(3 / 16), q, 0.1).τ_asymmetric == \
        pytest.approx(0.5 + (3 / 16) / 0.3)
    m = L.collision_model_from_viscosity(L.MRT, q, 0.1)
    assert len(m.τs) == 5 and m.τs[0] == pytest.approx(0.8)


@pytest.mark.parametrize("qname,n", [("D2Q4", 2), ("D2Q5", 2), ("D2Q9", 2), ("D2Q13", 2),
                                     ("D2Q17", 4), ("D2Q21", 4), ("D2Q37", 4)])
def test_mrt_relaxation_count(qname, n):
    """MRT(q, tau) fills round(order/2) times, Julia's ties-to-even (mrt.jl:19-22)."""
    import lbmhip as L
    q = L.Quadratures[qname]()
    assert len(L.MRT(q, 0.8).τs) == n
    assert L.MRT(q, 0.7, 0.9).τs == ([0.7, 0.9] * n)[:n]


def test_boundary_conditions():
    import lbmhip as L
    b = L.BounceBack(L.North(), range(1, 11), range(1, 6))
    assert b.xs == (1, 10) and b.ys == (1, 5)
    abi = b.to_abi()
    assert (abi.kind, abi.direction, abi.x0, abi.x1, abi.y0, abi.y1) == (0, 0, 1, 10, 1, 5)
    w = L.MovingWall(L.North(), (1, 3), (1, 4), [0.05, 0])
    assert w.u == (0.05, 0.0) and w.ρ == 1.0
    with pytest.raises(NotImplementedError):
        L.MovingWall(L.South(), (1, 3), (1, 4), [0.05, 0])
    with pytest.raises(ValueError):
        L.BounceBack(L.East(), range(1, 10, 2), (1, 2))


def test_opposite_is_one_based():
    import lbmhip as L
    q = L.D2Q9()
    assert [L.opposite(q, i) for i in range(1, 10)] == [1, 6, 7, 8, 9, 2, 3, 4, 5]
    with pytest.raises(IndexError):
        L.opposite(q, 0)


@pytest.mark.parametrize("qname", ["D2Q4", "D2Q5", "D2Q9", "D2Q13", "D2Q17", "D2Q21", "D2Q37"])
def test_initial_populations_match_oracle(oracle_lib, qname):
    """The host initial condition is the collide equilibrium to rounding."""
    import lbmhip as L
    q = L.Quadratures[qname]()
    rng = np.random.default_rng(3)
    rho = 1 + 0.01 * rng.standard_normal((3, 2))
    ux, uy = 0.03 * rng.standard_normal((2, 3, 2))
    f = L.equilibrium_populations(q, rho, ux, uy)
    for x in range(3):
        for y in range(2):
            ref = oracle_lib.equilibrium(q.id, rho[x, y], np.array([ux[x, y], uy[x, y]]))
            np.testing.assert_allclose(f[x, y], ref, rtol=0, atol=1e-16)


def test_cell_centres():
    import lbmhip as L
    np.testing.assert_allclose(L.cell_centres(4, 1.0), [0.125, 0.375, 0.625, 0.875])


def test_bench_refuses_multi_gpu():
    import subprocess
    import sys
    import os
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(repo, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_average_clock_pairs_samples_per_compute_unit():
    # Two compute units whose shader-clock counters start far apart: pairing
    # by location key gives each unit's own rate; pairing across units would
    # give nonsense (here a negative rate).
    import numpy as np

    import lbmhip as L

    tick = 100.0
    a = np.array([[0x0103, 5_000_000_000, 1000], [0x0204, 10, 1000]], dtype=np.uint64)
    b = np.array([[0x0204, 10 + 2_400_000, 1100], [0x0103, 5_000_000_000 + 2_000_000, 1100],
                  [0x0305, 7, 1100]], dtype=np.uint64)
    mhz = L.average_clock_mhz((a, tick), (b, tick))
    assert mhz == np.median([2_400_000 / 100 * tick, 2_000_000 / 100 * tick])
    assert L.average_clock_mhz((a[:1], tick), (b[2:], tick)) is None
