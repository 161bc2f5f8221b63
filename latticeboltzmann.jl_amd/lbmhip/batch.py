"""Batches of independent periodic lattices advanced together.

Every time step of every lattice is one persistent kernel launch over a tile
table; runs of fused steps replay a captured hipGraph.  Each lattice follows
simulate(model, time) (src/lattice_boltzmann_model.jl:60-77) exactly as a
single-lattice model does; the batch has no counterpart in the reference.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .quadratures import Quadrature


class LatticeBatch:
    def __init__(self, q: Quadrature, sizes: Sequence, tau_s: Sequence[float],
                 tau_a: Optional[Sequence[float]] = None, storage: str = "f32",
                 rho0: float = 1.0):
        n = len(sizes)
        nx = (ctypes.c_int * n)(*[int(s[0]) for s in sizes])
        ny = (ctypes.c_int * n)(*[int(s[1]) for s in sizes])
        ts = np.ascontiguousarray(tau_s, dtype=np.float64)
        ta = None if tau_a is None else np.ascontiguousarray(tau_a, dtype=np.float64)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().lbm_batch_create(ctypes.byref(h), q.id, _lib.STORAGES[storage], n,
                                               nx, ny, _lib.dptr(ts), _lib.dptr(ta), float(rho0)),
                   "lbm_batch_create")
        self._h = h
        self.quadrature = q
        self.sizes = [(int(s[0]), int(s[1])) for s in sizes]
        self.storage = storage

    def __len__(self):
        return len(self.sizes)

    @property
    def nodes(self) -> int:
        return int(sum(a * b for a, b in self.sizes))

    def upload(self, k: int, f: np.ndarray) -> None:
        a = np.asfortranarray(f, dtype=np.float64)
        if a.shape != self.sizes[k] + (self.quadrature.Q,):
            raise ValueError(f"lattice {k} is {self.sizes[k]}, got {a.shape}")
        _lib.check(_lib.lib().lbm_batch_upload(self._h, int(k), _lib.dptr(a)), "lbm_batch_upload")

    def download(self, k: int, role: str = "f_stream") -> np.ndarray:
        out = np.empty(self.sizes[k] + (self.quadrature.Q,), order="F")
        r = {"f_stream": 0, "f_collision": 1}[role]
        _lib.check(_lib.lib().lbm_batch_download(self._h, int(k), r, _lib.dptr(out)),
                   "lbm_batch_download")
        return out

    def set_launch(self, block: int = 0, grid: int = 0, nodes_per_thread: int = None) -> None:
        _lib.check(_lib.lib().lbm_batch_set_launch(self._h, int(block), int(grid)),
                   "lbm_batch_set_launch")
        if nodes_per_thread is not None:
            _lib.check(_lib.lib().lbm_batch_set_nodes_per_thread(self._h, int(nodes_per_thread)),
                       "lbm_batch_set_nodes_per_thread")

    def launch(self) -> dict:
        """The launch shape in use: threads per block, grid (0 = one block per
        tile) and nodes per thread."""
        b, g, v = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.lib().lbm_batch_get_launch(self._h, ctypes.byref(b), ctypes.byref(g),
                                                   ctypes.byref(v)), "lbm_batch_get_launch")
        return {"block": b.value, "grid": g.value, "nodes_per_thread": v.value}

    def autotune(self, trial_steps: int = 5) -> dict:
        b, g = ctypes.c_int(), ctypes.c_int()
        us = ctypes.c_double()
        _lib.check(_lib.lib().lbm_batch_autotune(self._h, int(trial_steps), ctypes.byref(b),
                                                 ctypes.byref(g), ctypes.byref(us)),
                   "lbm_batch_autotune")
        out = self.launch()
        out["us_per_step"] = us.value
        return out

    def simulate(self, nsteps: int, graph_chunk: int = 100) -> float:
        """Advances every lattice nsteps; returns the milliseconds of the fused steps."""
        ms = ctypes.c_double()
        _lib.check(_lib.lib().lbm_batch_simulate(self._h, int(nsteps), int(graph_chunk),
                                                 ctypes.byref(ms)), "lbm_batch_simulate")
        return ms.value

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.lib().lbm_batch_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
