"""Pins the oracle's velocity sets (quadrature constants and opposite maps).

Restates test/quadrature.jl:1-19 (sum w = 1, sum_i w_i H_n(c_i) = 0 for
n <= order/2), test/quadrature.jl:124-130 (c_i + c_opp(i) = 0), and checks
the C restatement against the independently transcribed NumPy tables.
"""
import numpy as np
import pytest

from lbm_cases import QIDS
from oracle import lbm_numpy as ln


@pytest.mark.parametrize("qid", QIDS)
def test_tables_agree_between_restatements(oracle_lib, qid):
    c, w, s, order, opp, feq_order = oracle_lib.quad_arrays(qid)
    q = ln.by_id(qid)
    assert np.array_equal(c, q.c)
    assert np.array_equal(w, q.w)  # bit-identical doubles
    assert s == q.s
    assert order == q.order and feq_order == q.feq_order
    assert np.array_equal(opp, q.opposite())


@pytest.mark.parametrize("qid", QIDS)
def test_weights_sum_to_one(oracle_lib, qid):
    _, w, *_ = oracle_lib.quad_arrays(qid)
    assert abs(w.sum() - 1.0) < 1e-14


@pytest.mark.parametrize("qid", QIDS)
def test_hermite_symmetry(oracle_lib, qid):
    """sum_i w_i H_n(c_i) = 0 for 1 <= n <= div(order, 2) (atol 1e-15).  With
    n = 2 this is sum_i w_i c_a c_b = delta_ab / s, which pins s."""
    c, w, s, order, *_ = oracle_lib.quad_arrays(qid)
    cs = 1.0 / s
    for n in range(1, order // 2 + 1):
        H = sum(w[i] * ln.hermite_tensor(n, c[:, i], cs) for i in range(w.size))
        assert np.all(np.abs(H) <= 1e-15), (qid, n, np.abs(H).max())


@pytest.mark.parametrize("qid", QIDS)
def test_opposite(oracle_lib, qid):
    c, _, _, _, opp, _ = oracle_lib.quad_arrays(qid)
    for i in range(c.shape[1]):
        assert np.all(c[:, i] + c[:, opp[i]] == 0)


def test_d2q9_known_values(oracle_lib):
    """D2Q9.jl:20-38 spelled out."""
    c, w, s, order, opp, _ = oracle_lib.quad_arrays(9)
    assert c.tolist() == [[0, -1, -1, -1, 0, 1, 1, 1, 0], [0, 1, 0, -1, -1, -1, 0, 1, 1]]
    assert w.tolist() == [4 / 9, 1 / 36, 1 / 9, 1 / 36, 1 / 9, 1 / 36, 1 / 9, 1 / 36, 1 / 9]
    assert s == 3.0 and order == 5
    assert (opp + 1).tolist() == [1, 6, 7, 8, 9, 2, 3, 4, 5]


def test_d2q21_zero_weight_planes(oracle_lib):
    """D2Q21 stores 25 populations; the last four (+-3, +-3) have weight 0."""
    c, w, *_ = oracle_lib.quad_arrays(21)
    assert w.size == 25 and np.all(w[21:] == 0.0) and np.all(w[:21] > 0.0)
    assert np.all(np.abs(c[:, 21:]) == 3)


def test_multispeed_isotropy_pins_s(oracle_lib):
    """A plausible transcription mistake in s (e.g. D2Q21's overridden
    cs = sqrt(2/3) of D2Q21.jl:16) fails the second-order isotropy check."""
    c, w, s, *_ = oracle_lib.quad_arrays(21)
    second = np.einsum("i,ai,bi->ab", w, c.astype(float), c.astype(float))
    np.testing.assert_allclose(second, np.eye(2) / s, atol=1e-15)
    wrong = 1.0 / np.sqrt(2.0 / 3.0)
    assert not np.allclose(second, np.eye(2) / wrong, atol=1e-6)
