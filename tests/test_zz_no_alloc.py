"""lbm_reduce_noalloc (the allocation-free variant of lbm_reduce, whose
partial sums live in a static device array) through the C ABI against
lbm_reduce: the same five sums, bit for bit, on two calls in a row -- the
second on a smaller lattice, after the first left its partials in the shared
scratch -- and on a lattice large enough to use every partial row."""
import ctypes

import numpy as np
import pytest

from lbm_cases import random_populations

pytestmark = pytest.mark.gpu


def _sums(fn, field):
    out = np.zeros(5)
    rc = fn(field.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return rc, out


@pytest.mark.parametrize("storage", ["f64", "f32", "f16dev", "bf16dev"])
def test_noalloc_matches_reduce(storage):
    import lbmhip as L
    import lbmhip._lib as lib
    C = lib.lib()
    q = L.D2Q9()
    for (nx, ny, seed) in ((300, 257, 1), (40, 9, 2), (1024, 600, 3)):
        f = L.HipLattice.from_numpy(q, random_populations(9, nx, ny, seed=seed), storage=storage)
        rc_a, a = _sums(C.lbm_reduce_noalloc, f)
        rc_b, b = _sums(C.lbm_reduce, f)
        assert rc_a == 0 and rc_b == 0, lib.lib().lbm_last_error()
        assert np.array_equal(a, b), (nx, ny, a, b)


def test_noalloc_two_calls_second_smaller():
    import lbmhip as L
    import lbmhip._lib as lib
    C = lib.lib()
    q = L.D2Q37()
    big = L.HipLattice.from_numpy(q, random_populations(37, 700, 500, seed=4))
    small = L.HipLattice.from_numpy(q, random_populations(37, 11, 7, seed=5))
    ref_big = _sums(C.lbm_reduce, big)[1]
    ref_small = _sums(C.lbm_reduce, small)[1]
    rc1, got_big = _sums(C.lbm_reduce_noalloc, big)
    rc2, got_small = _sums(C.lbm_reduce_noalloc, small)
    assert rc1 == 0 and rc2 == 0
    assert np.array_equal(got_big, ref_big)
    assert np.array_equal(got_small, ref_small)


def test_noalloc_rejects_neq_storage():
    import lbmhip as L
    import lbmhip._lib as lib
    C = lib.lib()
    f = L.HipLattice.from_numpy(L.D2Q9(), random_populations(9, 8, 8, seed=6), storage="f16neq")
    rc, _ = _sums(C.lbm_reduce_noalloc, f)
    assert rc != 0 and b"lbm_reduce" in C.lbm_last_error()
