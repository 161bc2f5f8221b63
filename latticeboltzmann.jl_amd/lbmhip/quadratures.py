"""Velocity sets D2Q4 ... D2Q37 (src/quadratures.jl, src/quadratures/D2Q*.jl).

The constants are read from liblbm_hip.so (``lbm_quadrature_info``) so that
host and kernels share one table.  Attribute names follow the reference:
``abscissae`` (2 x Q integers), ``weights``, ``speed_of_sound_squared``
(which, as in the reference, holds 1/c_s^2), ``order(q)`` and
``opposite(q, idx)`` with 1-based ``idx`` as in Julia.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class Quadrature:
    """Abstract velocity set (src/quadratures.jl:1)."""

    id: int = 0
    name: str = ""

    def __init__(self):
        L = _lib.lib()
        n = ctypes.c_int()
        cx = np.zeros(37)
        cy = np.zeros(37)
        w = np.zeros(37)
        s = ctypes.c_double()
        order = ctypes.c_int()
        opp = (ctypes.c_int * 37)()
        _lib.check(L.lbm_quadrature_info(self.id, ctypes.byref(n), _lib.dptr(cx), _lib.dptr(cy),
                                         _lib.dptr(w), ctypes.byref(s), ctypes.byref(order), opp),
                   "lbm_quadrature_info")
        Q = n.value
        self.abscissae = np.array([cx[:Q], cy[:Q]]).astype(np.int64)
        self.weights = w[:Q].copy()
        self.speed_of_sound_squared = s.value
        self._order = order.value
        self._opp = np.array(opp[:Q], dtype=np.int64)

    @property
    def Q(self) -> int:
        return self.weights.size

    def __repr__(self) -> str:
        return self.name

    def __eq__(self, other) -> bool:
        return isinstance(other, Quadrature) and other.id == self.id

    def __hash__(self) -> int:
        return hash(self.id)


def _make(qid: int):
    cls = type(f"D2Q{qid}", (Quadrature,), {"id": qid, "name": f"D2Q{qid}"})
    return cls


D2Q4 = _make(4)
D2Q5 = _make(5)
D2Q9 = _make(9)
D2Q13 = _make(13)
D2Q17 = _make(17)
D2Q21 = _make(21)
D2Q37 = _make(37)

Quadratures = {"D2Q4": D2Q4, "D2Q5": D2Q5, "D2Q9": D2Q9, "D2Q13": D2Q13, "D2Q17": D2Q17,
               "D2Q21": D2Q21, "D2Q37": D2Q37}


def order(q: Quadrature) -> int:
    """order(q) -- quadratures/D2Q*.jl"""
    return q._order


def opposite(q: Quadrature, idx: int) -> int:
    """opposite(q, idx) with 1-based idx (quadratures.jl:11-19, D2Q4.jl:26-31,
    D2Q5.jl:34-42, D2Q9.jl:30-38)."""
    if not 1 <= idx <= q.Q:
        raise IndexError(f"population index {idx} outside 1:{q.Q}")
    return int(q._opp[idx - 1]) + 1


def dimension(q: Quadrature) -> int:
    return 2
