"""GPU parity of the iterative initialization of Mei et al. and of the node
accessor, through the C ABI, against the oracle.  The FP64 path is bit-exact:
the nonlinear term is formed on the host in the oracle's order and the
collision kernel evaluates the reference's expression."""
import numpy as np
import pytest

from lbm_cases import QIDS, bc_dicts, random_populations, to_lbmhip

pytestmark = pytest.mark.gpu


def _L():
    import lbmhip
    return lbmhip


def _q(qid):
    return _L().Quadratures[f"D2Q{qid}"]()


def _tgv_u(nx, ny, u0=0.01):
    x = (np.arange(nx) + 0.5) * 2 * np.pi / nx
    y = (np.arange(ny) + 0.5) * 2 * np.pi / ny
    X, Y = np.meshgrid(x, y, indexing="ij")
    return np.asfortranarray(np.stack([u0 * np.cos(X) * np.sin(Y), -u0 * np.sin(X) * np.cos(Y)],
                                      axis=-1))


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("tau", [1.0, 0.8])
def test_collide_mei_bit_exact(oracle_lib, qid, tau):
    L = _L()
    nx, ny = 11, 7
    q = _q(qid)
    f = random_populations(qid, nx, ny, seed=qid)
    u = _tgv_u(nx, ny, 0.02)
    cm = L.IterativeInitializationCollisionModel(tau, u)
    a = L.HipLattice.from_numpy(q, f)
    b = L.HipLattice(q, nx, ny)
    L.collide_(cm, q, a, b)
    ref = np.zeros_like(f, order="F")
    oracle_lib.collide_mei(qid, f, ref, tau, oracle_lib.mei_nonlinear(qid, u))
    assert np.array_equal(b.to_numpy(), ref)


@pytest.mark.parametrize("qid", [9, 17, 37])
@pytest.mark.parametrize("bcname", ["periodic", "couette"])
def test_simulate_with_mei_model_bit_exact(oracle_lib, qid, bcname):
    L = _L()
    nx, ny, steps = 12, 9, 5
    q = _q(qid)
    f = random_populations(qid, nx, ny, seed=3)
    u = _tgv_u(nx, ny, 0.02)
    bcd = bc_dicts(bcname, nx, ny)
    m = L.LatticeBoltzmannModel.from_populations(q, f, L.IterativeInitializationCollisionModel(0.9, u),
                                                 to_lbmhip(bcd))
    L.simulate(m, range(steps))
    N = oracle_lib.mei_nonlinear(qid, u)
    S = np.asfortranarray(f.copy())
    C = np.zeros_like(S, order="F")
    for _ in range(steps):
        oracle_lib.collide_mei(qid, S, C, 0.9, N)
        oracle_lib.stream(qid, C, S)
        oracle_lib.apply_bcs(qid, bcd, S, C)
    assert np.array_equal(m.f_stream.to_numpy(), S)
    assert np.array_equal(m.f_collision.to_numpy(), C)


@pytest.mark.parametrize("eps", [1e-7, 1e-10])
def test_initialize_mei_matches_oracle(oracle_lib, eps):
    L = _L()
    pb = L.TaylorGreenVortex(scale=2)                 # 32 x 32
    q = L.D2Q9()
    f, steps = L.initialize_mei(L.IterativeInitializationMeiEtAl(1.0, eps), q, pb)
    ref, _, ref_steps = oracle_lib.mei_initialize(9, pb.lattice_velocity_field(), 1.0, eps)
    assert steps == ref_steps
    assert np.array_equal(f.to_numpy(), ref)


def test_initialize_dispatch():
    L = _L()
    pb = L.TaylorGreenVortex(scale=1)                 # 16 x 16
    q = L.D2Q9()
    f = L.initialize(L.AnalyticalEquilibrium(), q, pb)
    np.testing.assert_allclose(f.to_numpy(), pb.initial_populations(q), rtol=0, atol=2e-16)
    g = L.initialize(L.IterativeInitializationMeiEtAl(), q, pb)
    h, _ = L.initialize_mei(L.IterativeInitializationMeiEtAl(), q, pb)
    assert np.array_equal(g.to_numpy(), h.to_numpy())
    with pytest.raises(ValueError):
        L.initialize(L.IterativeInitializationMeiEtAl(), q, pb, storage="f32")


def test_initialize_mei_uniform_velocity_two_steps():
    L = _L()

    class Uniform:
        NX, NY = 40, 24

        def lattice_velocity_field(self):
            return np.broadcast_to([0.02, 0.005], (self.NX, self.NY, 2))

    q = L.D2Q37()
    f, steps = L.initialize_mei(L.IterativeInitializationMeiEtAl(), q, Uniform())
    assert steps == 2


def test_mei_errors():
    L = _L()
    q = L.D2Q9()
    u = _tgv_u(8, 6)
    cm = L.IterativeInitializationCollisionModel(1.0, u)
    a = L.HipLattice(q, 8, 6, storage="f32")
    b = L.HipLattice(q, 8, 6, storage="f32")
    with pytest.raises(L.LbmError):
        L.collide_(cm, q, a, b)                       # FP64 only
    c = L.HipLattice(q, 8, 7)
    d = L.HipLattice(q, 8, 7)
    with pytest.raises(L.LbmError):
        L.collide_(cm, q, c, d)                       # velocity field of another size


@pytest.mark.parametrize("storage", ["f64", "f32", "f16dev", "bf16dev", "f16neq", "bf16neq"])
@pytest.mark.parametrize("qid", [9, 21, 37])
def test_field_node_matches_download(qid, storage):
    L = _L()
    nx, ny = 13, 6
    q = _q(qid)
    lat = L.HipLattice.from_numpy(q, random_populations(qid, nx, ny, seed=5), storage=storage)
    full = lat.to_numpy()
    for (x, y) in [(1, 1), (nx, ny), (7, 3), (1, ny)]:
        assert np.array_equal(lat.node(x, y), full[x - 1, y - 1, :])
    with pytest.raises(L.LbmError):
        lat.node(nx + 1, 1)
