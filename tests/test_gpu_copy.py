"""The device copy behind ``copy(f)`` (lbm_field_copy; the model's
f_collision = copy(f_stream), lattice_boltzmann_model.jl:22-23) for every
storage, on lattices whose byte counts divide into 16-byte lanes (64 x 64
FP64 and FP32) and on ones that do not (odd sizes: the 8-byte and 2-byte
paths): the copy equals its source bit for bit, and copying into an existing
field leaves the source unchanged."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STORAGES = ["f64", "f32", "f16dev", "bf16dev", "f16neq", "bf16neq"]


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("qid,nx,ny", [(9, 64, 64), (9, 7, 5), (9, 1, 1), (17, 9, 7), (37, 33, 3)])
def test_field_copy_bit_exact(storage, qid, nx, ny):
    import lbmhip as L
    q = L.Quadratures[f"D2Q{qid}"]()
    rng = np.random.default_rng(1000 * qid + 10 * nx + ny)
    rho = 1.0 + 0.01 * rng.standard_normal((nx, ny))
    ux, uy = 0.02 * rng.standard_normal((2, nx, ny))
    f = L.equilibrium_populations(q, rho, ux, uy)
    print(f"copy D2Q{qid} {nx}x{ny} {storage}", flush=True)
    a = L.HipLattice.from_numpy(q, f, storage)
    ref = a.to_numpy()
    b = a.copy()
    assert np.array_equal(b.to_numpy(), ref)
    c = L.HipLattice(q, nx, ny, storage)
    c.copyto_(a)
    assert np.array_equal(c.to_numpy(), ref)
    assert np.array_equal(a.to_numpy(), ref)
