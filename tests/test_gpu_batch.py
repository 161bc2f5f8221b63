"""Batched independent lattices: each lattice of a batch evolves exactly as
it would alone (FP64: bit for bit against the oracle's step loop)."""
import numpy as np
import pytest

from lbm_cases import random_populations

pytestmark = pytest.mark.gpu

SIZES = [(5, 7), (16, 16), (33, 20), (1, 4), (64, 3), (40, 41), (2, 2), (70, 9)]


def _batch(storage, tau_a=True):
    import lbmhip as L
    rng = np.random.default_rng(0)
    ts = rng.uniform(0.52, 1.5, len(SIZES))
    ta = (0.5 + 0.25 / (ts - 0.5)) if tau_a else None
    b = L.LatticeBatch(L.D2Q9(), SIZES, ts, ta, storage=storage)
    f0 = [random_populations(9, nx, ny, seed=k) for k, (nx, ny) in enumerate(SIZES)]
    for k, f in enumerate(f0):
        b.upload(k, f)
    return b, f0, ts, ta


@pytest.mark.parametrize("chunk", [0, 4])
@pytest.mark.parametrize("trt", [False, True])
def test_batch_f64_bit_exact(oracle_lib, chunk, trt):
    b, f0, ts, ta = _batch("f64", trt)
    b.simulate(13, graph_chunk=chunk)
    for k, f in enumerate(f0):
        S, C = f.copy(order="F"), f.copy(order="F")
        taus = [ts[k], ta[k]] if trt else [ts[k]]
        oracle_lib.simulate(9, S, C, 1 if trt else 0, taus, 13)
        assert np.array_equal(b.download(k, "f_stream"), S), k
        assert np.array_equal(b.download(k, "f_collision"), C), k


@pytest.mark.parametrize("nodes", [1, 2, 4])
@pytest.mark.parametrize("block,grid", [(64, 0), (256, 0), (128, 3)])
def test_batch_launch_shapes_bit_exact(oracle_lib, nodes, block, grid):
    """every tile shape (block x nodes per thread), one block per tile or a
    3-block grid striding over the table, gives the oracle's bits; lattices
    smaller than one tile leave lanes that load a clamped node and store
    nothing"""
    b, f0, ts, ta = _batch("f64")
    b.set_launch(block, grid, nodes_per_thread=nodes)
    assert b.launch() == {"block": block, "grid": grid, "nodes_per_thread": nodes}
    b.simulate(7, graph_chunk=0)
    b.simulate(6, graph_chunk=3)
    for k, f in enumerate(f0):
        S, C = f.copy(order="F"), f.copy(order="F")
        oracle_lib.simulate(9, S, C, 1, [ts[k], ta[k]], 13)
        assert np.array_equal(b.download(k, "f_stream"), S), k
        assert np.array_equal(b.download(k, "f_collision"), C), k


def test_batch_launch_errors():
    import lbmhip as L
    b, _, _, _ = _batch("f32")
    with pytest.raises(L.LbmError):
        b.set_launch(96, 0)
    with pytest.raises(L.LbmError):
        b.set_launch(128, 0, nodes_per_thread=3)
    assert b.launch()["nodes_per_thread"] == 1


def test_batch_continues_across_calls(oracle_lib):
    b, f0, ts, ta = _batch("f64")
    b.simulate(5, graph_chunk=2)
    b.simulate(8, graph_chunk=4)
    for k, f in enumerate(f0):
        S, C = f.copy(order="F"), f.copy(order="F")
        oracle_lib.simulate(9, S, C, 1, [ts[k], ta[k]], 5)
        oracle_lib.simulate(9, S, C, 1, [ts[k], ta[k]], 8)
        assert np.array_equal(b.download(k), S), k


def test_batch_f32_graph_equals_direct():
    a, f0, _, _ = _batch("f32")
    c, _, _, _ = _batch("f32")
    a.simulate(41, graph_chunk=10)
    c.simulate(41, graph_chunk=0)
    for k in range(len(SIZES)):
        assert np.array_equal(a.download(k), c.download(k)), k


def test_batch_f32_tracks_single_lattice():
    import lbmhip as L
    b, f0, ts, ta = _batch("f32")
    b.simulate(20)
    for k, f in enumerate(f0):
        m = L.LatticeBoltzmannModel.from_populations(L.D2Q9(), f, L.TRT.fields(ts[k], ta[k]),
                                                     storage="f32")
        L.simulate(m, range(20))
        np.testing.assert_allclose(b.download(k), m.f_stream.to_numpy(), rtol=0, atol=2e-7)


def test_batch_autotune_keeps_state(oracle_lib):
    b, f0, ts, ta = _batch("f64")
    best = b.autotune(2)
    assert best["block"] in (64, 128, 256) and best["grid"] > 0
    assert best["nodes_per_thread"] in (1, 2, 4)
    for k, f in enumerate(f0):
        assert np.array_equal(b.download(k), f), k
    b.simulate(3)
    S, C = f0[2].copy(order="F"), f0[2].copy(order="F")
    oracle_lib.simulate(9, S, C, 1, [ts[2], ta[2]], 3)
    assert np.array_equal(b.download(2), S)


def test_batch_rejects_other_quadratures():
    import lbmhip as L
    with pytest.raises(L.LbmError, match="D2Q9"):
        L.LatticeBatch(L.D2Q17(), [(4, 4)], [0.8])


def test_batch_graph_after_tile_table_rebuilt(oracle_lib):
    """A captured step graph holds the tile table's address: changing the
    nodes per thread rebuilds the table, which must drop the graph, including
    a later switch back to the shape the graph was captured with."""
    b, f0, ts, ta = _batch("f64")
    b.set_launch(128, 0, nodes_per_thread=1)
    b.simulate(6, graph_chunk=6)
    b.set_launch(128, 0, nodes_per_thread=2)
    b.simulate(6, graph_chunk=6)
    b.set_launch(128, 0, nodes_per_thread=1)
    b.simulate(6, graph_chunk=6)
    for k, f in enumerate(f0):
        S, C = f.copy(order="F"), f.copy(order="F")
        oracle_lib.simulate(9, S, C, 1, [ts[k], ta[k]], 18)
        assert np.array_equal(b.download(k, "f_stream"), S), k
        assert np.array_equal(b.download(k, "f_collision"), C), k
