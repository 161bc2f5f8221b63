"""Pins the oracle's streaming and boundary conditions.

* stream! is a per-plane periodic permutation: equal to numpy.roll by c_i,
  for every lattice size including |c| >= n (test/quadrature.jl:135-147);
* collide(tau = 1) + stream on a 1x1 equilibrium is the identity
  (test/quadrature.jl:98-121);
* halfway bounce-back on a one-row lattice maps i -> opp(i);
* ordered BC lists (last writer wins) against the NumPy mask formulation;
* the moving-wall term 2 w_i s rho_w u_w . c_i (moving_wall.jl:17-38).
"""
import numpy as np
import pytest

from lbm_cases import BC_NAMES, QIDS, bc_dicts, random_populations
from oracle import lbm_numpy as ln


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("shape", [(7, 5), (1, 1), (2, 3), (1, 5), (4, 1)])
def test_stream_is_roll(oracle_lib, qid, shape):
    q = ln.by_id(qid)
    f = random_populations(qid, *shape, seed=7)
    out = np.zeros_like(f, order="F")
    oracle_lib.stream(qid, f, out)
    for i in range(q.Q):
        ref = np.roll(f[:, :, i], shift=(int(q.c[0, i]), int(q.c[1, i])), axis=(0, 1))
        assert np.array_equal(out[:, :, i], ref)


@pytest.mark.parametrize("qid", QIDS)
def test_multispeed_1x1_stream_identity(oracle_lib, qid):
    f = np.asfortranarray(oracle_lib.equilibrium(qid, 1.0, np.zeros(2)).reshape(1, 1, -1))
    out = np.zeros_like(f, order="F")
    oracle_lib.stream(qid, f, out)
    assert np.array_equal(f, out)


@pytest.mark.parametrize("qid", QIDS)
def test_collide_stream_fixed_point_1x1(oracle_lib, qid):
    f = np.asfortranarray(oracle_lib.equilibrium(qid, 1.0, np.array([0.1, 0.1])).reshape(1, 1, -1))
    S, C = f.copy(order="F"), f.copy(order="F")
    oracle_lib.simulate(qid, S, C, 0, [1.0], 4)
    np.testing.assert_allclose(S, f, atol=1e-15)


@pytest.mark.parametrize("qid", QIDS)
def test_bounce_back_one_row(oracle_lib, qid):
    q = ln.by_id(qid)
    opp = q.opposite()
    nx = 5
    fold = random_populations(qid, nx, 1, seed=8)
    fnew = random_populations(qid, nx, 1, seed=9)
    got = fnew.copy(order="F")
    oracle_lib.apply_bcs(qid, bc_dicts("poiseuille", nx, 1), got, fold)
    for i in range(q.Q):
        if q.c[1, i] != 0:
            assert np.array_equal(got[:, 0, i], fold[:, 0, opp[i]])
        else:
            assert np.array_equal(got[:, 0, i], fnew[:, 0, i])


@pytest.mark.parametrize("qid", QIDS)
@pytest.mark.parametrize("bcname", BC_NAMES[1:])
@pytest.mark.parametrize("shape", [(9, 8), (3, 5), (1, 4)])
def test_bcs_match_numpy(oracle_lib, qid, bcname, shape):
    q = ln.by_id(qid)
    nx, ny = shape
    fold = random_populations(qid, nx, ny, seed=10)
    fnew = random_populations(qid, nx, ny, seed=11)
    bcs = bc_dicts(bcname, nx, ny)
    got = fnew.copy(order="F")
    oracle_lib.apply_bcs(qid, bcs, got, fold)
    np.testing.assert_allclose(got, ln.apply_bcs(q, bcs, fnew, fold), rtol=0, atol=1e-16)


def test_list_order_last_writer_wins(oracle_lib):
    """Two BCs hitting the same (node, population): the later one's value."""
    nx, ny = 4, 3
    fold = random_populations(9, nx, ny, seed=12)
    fnew = random_populations(9, nx, ny, seed=13)
    wall = dict(kind=1, direction=0, xs=(1, nx), ys=(1, ny), u=(0.05, 0.0), rho=1.0)
    bb = dict(kind=0, direction=0, xs=(1, nx), ys=(1, ny))
    a, b = fnew.copy(order="F"), fnew.copy(order="F")
    oracle_lib.apply_bcs(9, [wall, bb], a, fold)   # bounce-back last
    oracle_lib.apply_bcs(9, [bb, wall], b, fold)   # moving wall last
    q = ln.by_id(9)
    opp = q.opposite()
    for i in range(9):
        if q.c[1, i] < 0:  # pulled from beyond the north wall
            assert np.array_equal(a[:, ny - 1, i], fold[:, ny - 1, opp[i]])
            term = 2.0 * q.w[i] * q.s * 0.05 * q.c[0, i]
            np.testing.assert_allclose(b[:, ny - 1, i], fold[:, ny - 1, opp[i]] + term,
                                       rtol=0, atol=1e-17)


@pytest.mark.parametrize("qid", [9, 17, 37])
def test_moving_wall_momentum(oracle_lib, qid):
    """Replacing bounce-back by the moving wall adds exactly
    sum_i c_i 2 w_i s rho_w (u_w . c_i) over the replaced populations."""
    q = ln.by_id(qid)
    nx, ny = 6, 7
    fold = random_populations(qid, nx, ny, seed=14)
    fnew = random_populations(qid, nx, ny, seed=15)
    u = (0.04, 0.0)
    a, b = fnew.copy(order="F"), fnew.copy(order="F")
    oracle_lib.apply_bcs(qid, [dict(kind=0, direction=0, xs=(1, nx), ys=(1, ny))], a, fold)
    oracle_lib.apply_bcs(qid, [dict(kind=1, direction=0, xs=(1, nx), ys=(1, ny), u=u, rho=1.0)],
                         b, fold)
    d = b - a
    Y = np.arange(1, ny + 1)
    for i in range(q.Q):
        hit = (Y - q.c[1, i]) > ny
        expect = 2.0 * q.w[i] * q.s * (u[0] * q.c[0, i] + u[1] * q.c[1, i])
        np.testing.assert_allclose(d[:, hit, i], expect, rtol=1e-12, atol=1e-17)
        assert np.all(d[:, ~hit, i] == 0.0)


def test_out_of_range_bc_is_an_error(oracle_lib):
    f = random_populations(9, 3, 4, seed=1)
    g = f.copy(order="F")
    with pytest.raises(IndexError):
        oracle_lib.apply_bcs(9, [dict(kind=0, direction=0, xs=(1, 4), ys=(1, 4))], g, f)
    assert np.array_equal(g, f)
