"""The analytic initial conditions on the device (lbm_initialize_analytical)
against the oracle's restatement (lbo_initialize), through the C ABI.

FP64 storage is bit-exact for every quadrature and strategy: the Hermite-sum
equilibrium with temperature terms (AnalyticalEquilibrium,
analytical_equilibrium.jl:8-17; hermite.jl:10-77), and the equilibrium +
off-equilibrium starts (analytical_offequilibrium.jl:10-87).  Reduced
storages hold the oracle's FP64 populations converted as an upload converts
them (bit-identical to uploading the oracle's result).  The problem-level
dispatch (initialize with TaylorGreenVortex / the convergence study's TGV) is
checked against the same restatement, and the off-equilibrium start of the
TGV is checked against its closed form: its deviatoric stress is the
Newtonian -rho nu (grad u + grad u^T).
"""
import numpy as np
import pytest

from lbm_cases import QIDS

pytestmark = pytest.mark.gpu


def _L():
    import lbmhip
    return lbmhip


def _fields(nx, ny, seed, with_T=True):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.02 * rng.standard_normal((nx, ny))
    ux, uy = 0.03 * rng.standard_normal((2, nx, ny))
    T = 1.0 + 0.01 * rng.standard_normal((nx, ny)) if with_T else None
    G = 1e-3 * rng.standard_normal((nx, ny, 2, 2))
    return rho, ux, uy, T, G


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("with_T", [False, True])
def test_initialize_bit_exact(oracle_lib, qid, mode, with_T):
    L = _L()
    q = L.Quadratures[f"D2Q{qid}"]()
    nx, ny = 13, 7
    rho, ux, uy, T, G = _fields(nx, ny, qid + 10 * mode, with_T)
    grad = G if mode else None
    f = L.initialize_analytical(q, rho, ux, uy, T, grad, mode, tau=0.35, factor=1.5)
    ref = oracle_lib.initialize(qid, mode, rho, ux, uy, T=T, G=grad, tau=0.35, factor=1.5)
    got = f.to_numpy()
    assert np.array_equal(got, ref), np.abs(got - ref).max()


@pytest.mark.parametrize("storage", ["f32", "f16dev", "bf16dev", "f16neq", "bf16neq"])
@pytest.mark.parametrize("qid", [9, 37])
def test_initialize_reduced_storage(oracle_lib, storage, qid):
    L = _L()
    q = L.Quadratures[f"D2Q{qid}"]()
    nx, ny = 16, 5
    rho, ux, uy, T, G = _fields(nx, ny, 5)
    f = L.initialize_analytical(q, rho, ux, uy, T, G, 1, tau=0.2, storage=storage)
    ref = oracle_lib.initialize(qid, 1, rho, ux, uy, T=T, G=G, tau=0.2)
    up = L.HipLattice.from_numpy(q, ref, storage=storage)
    assert np.array_equal(f.to_numpy(), up.to_numpy())


def test_initialize_dispatch_taylor_green(oracle_lib):
    L = _L()
    q = L.D2Q9()
    pb = L.TaylorGreenVortex(scale=1)                 # 16 x 16
    f = L.initialize(L.AnalyticalEquilibrium(), q, pb)
    rho, ux, uy, T = pb.lattice_fields(q)
    ref = oracle_lib.initialize(9, 0, rho, ux, uy, T=T)
    assert np.array_equal(f.to_numpy(), ref)
    # the Hermite sum equals the closed form to rounding
    np.testing.assert_allclose(f.to_numpy(), pb.initial_populations(q), rtol=0, atol=2e-16)
    g = L.initialize(L.AnalyticalEquilibriumAndOffEquilibrium(), q, pb)
    mode, grad, factor = pb.offequilibrium_inputs()
    assert mode == 2
    ref = oracle_lib.initialize(9, 2, rho, ux, uy, T=T, G=grad,
                                tau=q.speed_of_sound_squared * pb.lattice_viscosity(),
                                factor=factor)
    assert np.array_equal(g.to_numpy(), ref)


@pytest.mark.parametrize("tau", [0.6, 1.0])
def test_tgv_offequilibrium_start_stress(oracle_lib, tau):
    """The convergence study's TGV (second_order_convergence.jl) started from
    equilibrium + off-equilibrium: bit-exact against the restatement, and
    its deviatoric stress (tau' = s nu, moments.jl:81-96) is -rho nu S."""
    L = _L()
    q = L.D2Q9()
    pb = L.TGV(q, tau, scale=4)                        # 64 x 64
    f = L.initialize(L.AnalyticalEquilibriumAndOffEquilibrium(), q, pb)
    rho, ux, uy, T = pb.lattice_fields(q)
    G = pb.velocity_gradient()
    ref = oracle_lib.initialize(9, 1, rho, ux, uy, T=T, G=G, tau=3.0 * pb.ν)
    assert np.array_equal(f.to_numpy(), ref)
    m = L.moments(f, 3.0 * pb.ν)
    S = G + np.swapaxes(G, -1, -2)
    tr = 0.5 * (S[..., 0, 0] + S[..., 1, 1])
    scale = np.abs(pb.ν * S).max()
    np.testing.assert_allclose(m.sigma_xx, -rho * pb.ν * (S[..., 0, 0] - tr), rtol=0,
                               atol=1e-9 * scale)
    np.testing.assert_allclose(m.sigma_xy, -rho * pb.ν * S[..., 0, 1], rtol=0, atol=1e-9 * scale)
    np.testing.assert_allclose(m.ux, ux, rtol=0, atol=1e-15)


def test_initialize_errors():
    L = _L()
    q = L.D2Q9()
    rho = np.ones((4, 4))
    with pytest.raises(L.LbmError):
        L.initialize_analytical(q, rho, 0.0, 0.0, strategy=1)   # off-equilibrium without G
    with pytest.raises(ValueError):
        L.initialize_analytical(q, rho, 0.0, 0.0, grad=np.zeros((4, 4, 2)), strategy=1)
