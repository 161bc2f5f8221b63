"""The FP32 weights and s of the reduced-precision storages
(``lbm_quadrature_f32``), on the CPU.

The FP32 collision conserves mass and momentum only as well as the FP32
constants satisfy the quadrature's identities sum_i w_i = 1 and
s sum_i w_i c_ix^2 = 1 (s = 1/c_s^2; the weights, abscissae and s of
src/quadratures/D2Q*.jl).  The checks, exact in float64 (sums of floats
times small integers):
* D2Q13/17/21/37: both identities hold to 1.5e-9 -- rounding each constant
  on its own leaves 8e-9 to 3e-8, so a library that hands out the plain
  rounding fails here;
* every constant sits within 3 ulps of its FP32 rounding, and populations of
  one weight group keep one weight (a tuned weight per population would break
  the lattice symmetries that make sum_i w_i c_i = 0);
* D2Q4/5/9 keep the plain rounding (s = 3 exactly on D2Q9).
"""
import ctypes

import numpy as np
import pytest

from lbmhip import _lib

QIDS = (4, 5, 9, 13, 17, 21, 37)


def _consts(qid):
    L = _lib.lib()
    n = ctypes.c_int()
    assert L.lbm_quadrature_info(qid, ctypes.byref(n), None, None, None, None, None, None) == 0
    Q = n.value
    cx, cy, w = (np.zeros(Q) for _ in range(3))
    s = ctypes.c_double()
    assert L.lbm_quadrature_info(qid, None, _lib.dptr(cx), _lib.dptr(cy), _lib.dptr(w),
                                 ctypes.byref(s), None, None) == 0
    w32 = np.zeros(Q, dtype=np.float32)
    s32 = ctypes.c_float()
    assert L.lbm_quadrature_f32(qid, w32.ctypes.data_as(ctypes.c_void_p),
                                ctypes.cast(ctypes.byref(s32), ctypes.c_void_p)) == 0
    return cx, cy, w, s.value, w32, np.float32(s32.value)


def _defects(cx, w32, s32):
    wd = w32.astype(np.float64)
    return abs(wd.sum() - 1.0), abs(float(s32) * float(np.sum(wd * cx * cx)) - 1.0)


def _ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


@pytest.mark.parametrize("qid", QIDS)
def test_f32_constants_near_rounding_and_symmetric(qid):
    cx, cy, w, s, w32, s32 = _consts(qid)
    nz = w != 0
    assert _ulps(w32[nz], w[nz].astype(np.float32)).max() <= 3
    assert _ulps(s32, np.float32(s)) <= 3
    for v in np.unique(w):
        assert len(set(w32[w == v].tolist())) == 1, f"D2Q{qid}: weight group {v} split"


@pytest.mark.parametrize("qid", (13, 17, 21, 37))
def test_f32_sum_rules_hold(qid):
    cx, cy, w, s, w32, s32 = _consts(qid)
    mass, mom = _defects(cx, w32, s32)
    plain = _defects(cx, w.astype(np.float32), np.float32(s))
    assert max(mass, mom) <= 1.5e-9, (qid, mass, mom)
    # the plain rounding is what the test rules out
    assert max(plain) > 4 * max(mass, mom), (qid, plain, mass, mom)
    # y moments are the same by symmetry
    assert abs(float(s32) * float(np.sum(w32.astype(np.float64) * cy * cy)) - 1.0) <= 1.5e-9


@pytest.mark.parametrize("qid", (4, 5, 9))
def test_f32_plain_rounding_on_small_sets(qid):
    cx, cy, w, s, w32, s32 = _consts(qid)
    assert np.array_equal(w32, w.astype(np.float32))
    assert s32 == np.float32(s)
