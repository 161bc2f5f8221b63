"""The C ABI library loads without a GPU, exports every entry point that
include/lbm_hip.h declares, holds the same quadrature constants as the
oracle (bit for bit), and reports errors instead of running on a CPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from lbm_cases import QIDS

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "lbm_hip.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(lbm_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "lbm_simulate" in names and "lbm_collide" in names and len(names) >= 25


def test_library_exports_every_declared_symbol():
    import lbmhip._lib as lib
    L = lib.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(lib.EXPORTS) == _declared()


def test_version_and_device_count_without_gpu():
    import lbmhip._lib as lib
    L = lib.lib()
    assert L.lbm_version() >= 10000
    assert L.lbm_device_count() >= 0


@pytest.mark.parametrize("qid", QIDS)
def test_quadrature_constants_match_oracle(oracle_lib, qid):
    import lbmhip
    q = lbmhip.Quadratures[f"D2Q{qid}"]()
    c, w, s, order, opp, _ = oracle_lib.quad_arrays(qid)
    assert np.array_equal(q.abscissae, c)
    assert np.array_equal(q.weights, w)
    assert q.speed_of_sound_squared == s and lbmhip.order(q) == order
    assert [lbmhip.opposite(q, i) - 1 for i in range(1, q.Q + 1)] == list(opp)


def test_unknown_quadrature_is_an_error():
    import lbmhip._lib as lib
    L = lib.lib()
    n = ctypes.c_int()
    assert L.lbm_quadrature_info(10, ctypes.byref(n), None, None, None, None, None, None) < 0
    assert b"D2Q10" in L.lbm_last_error()


def test_no_device_is_reported_not_emulated():
    import lbmhip
    if lbmhip.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(lbmhip.LbmError, match="no HIP device"):
        lbmhip.HipLattice(lbmhip.D2Q9(), 4, 4)
