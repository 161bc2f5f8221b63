"""Device population arrays.

``HipLattice`` is the device counterpart of the reference's
``Array{Float64,3}`` of size ``(nx, ny, Q)`` (src/initial_conditions.jl:13):
one contiguous HBM plane per discrete velocity, x fastest.  The operators in
:mod:`lbmhip.operators` take and mutate these arrays exactly where the
reference takes Julia arrays.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .quadratures import Quadrature


class HipLattice:
    """Populations f(x, y, i) on the GPU, in one of the storages
    ``"f64"`` (plain), ``"f32"``, ``"f16dev"``, ``"bf16dev"`` (deviations
    f_i - w_i rho0 in 16 bits; FP32 collision arithmetic)."""

    def __init__(self, q: Quadrature, nx: int, ny: int, storage: str = "f64", rho0: float = 1.0):
        if storage not in _lib.STORAGES:
            raise ValueError(f"unknown storage {storage!r}; one of {sorted(_lib.STORAGES)}")
        self.quadrature = q
        self.nx, self.ny = int(nx), int(ny)
        self.storage = storage
        self.rho0 = float(rho0)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().lbm_field_create(ctypes.byref(h), q.id, self.nx, self.ny,
                                               _lib.STORAGES[storage], self.rho0),
                   "lbm_field_create")
        self._h = h

    # -- Julia-array-like surface -------------------------------------------
    @property
    def shape(self):
        return (self.nx, self.ny, self.quadrature.Q)

    def size(self, dim=None):
        return self.shape if dim is None else self.shape[dim - 1]

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def upload(self, f: np.ndarray) -> "HipLattice":
        a = np.asarray(f, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"array shape {a.shape} does not match lattice {self.shape}")
        a = np.asfortranarray(a)
        _lib.check(_lib.lib().lbm_field_upload(self._h, _lib.dptr(a)), "lbm_field_upload")
        return self

    def node(self, x: int, y: int) -> np.ndarray:
        """f[x, y, :] with 1-based (x, y), read from the device in FP64"""
        out = np.empty(self.quadrature.Q)
        _lib.check(_lib.lib().lbm_field_node(self._h, int(x), int(y), _lib.dptr(out)),
                   "lbm_field_node")
        return out

    def to_numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64, order="F")
        _lib.check(_lib.lib().lbm_field_download(self._h, _lib.dptr(out)), "lbm_field_download")
        return out

    @classmethod
    def from_numpy(cls, q: Quadrature, f: np.ndarray, storage: str = "f64",
                   rho0: float = 1.0) -> "HipLattice":
        nx, ny, Q = np.shape(f)
        if Q != q.Q:
            raise ValueError(f"{Q} populations given for {q!r} (needs {q.Q})")
        return cls(q, nx, ny, storage, rho0).upload(f)

    def copy(self) -> "HipLattice":
        other = HipLattice(self.quadrature, self.nx, self.ny, self.storage, self.rho0)
        _lib.check(_lib.lib().lbm_field_copy(other._h, self._h), "lbm_field_copy")
        return other

    similar = copy

    def copyto_(self, src: "HipLattice") -> "HipLattice":
        _lib.check(_lib.lib().lbm_field_copy(self._h, src._h), "lbm_field_copy")
        return self

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.lib().lbm_field_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __repr__(self) -> str:
        return f"HipLattice({self.quadrature!r}, {self.nx}x{self.ny}, {self.storage})"
