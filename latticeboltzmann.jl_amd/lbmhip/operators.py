"""Step operators with the reference's call signatures.

Julia's mutating ``collide!`` / ``stream!`` / ``apply!`` become ``collide_`` /
``stream_`` / ``apply_`` (the trailing underscore marks in-place mutation).
Both the positional and the keyword forms of the reference are accepted:

* ``collide_(cm, q, f_in, f_out, time=0.0)`` and
  ``collide_(cm, q, f_old=..., f_new=..., time=...)`` (srt.jl:18-24,
  trt.jl:42-48, mrt.jl:49, collision_models.jl:19);
* ``stream_(q, f, f_new)`` and ``stream_(q, f_new=..., f_old=...)`` (stream.jl:18-19);
* ``apply_(bcs, q, f_new, f_old, time=0.0)`` (boundary_conditions.jl:6-20).

Each call launches the matching gfx950 kernel through liblbm_hip.so.
"""
from __future__ import annotations

import ctypes
from typing import Any, Sequence

import numpy as np

from . import _lib
from .boundary_conditions import BoundaryCondition
from .collision_models import MRT, SRT, TRT, CollisionModel, IterativeInitializationCollisionModel
from .lattice import HipLattice
from .quadratures import Quadrature


class NativeModel:
    """An ``lbm_model`` handle (collision model + force + BC list) that is
    reconfigured only when the requested configuration changes."""

    def __init__(self, q: Quadrature):
        self.q = q
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().lbm_model_create(ctypes.byref(h), q.id), "lbm_model_create")
        self._h = h
        self._coll_key = None
        self._force_key = None
        self._bcs_key = None
        self._launch_key = None

    @property
    def handle(self):
        return self._h

    def set_collision(self, cm: CollisionModel) -> None:
        L = _lib.lib()
        if isinstance(cm, SRT):
            key = ("srt", float(cm.τ))
            if key != self._coll_key:
                _lib.check(L.lbm_model_set_srt(self._h, key[1]), "lbm_model_set_srt")
        elif isinstance(cm, TRT):
            key = ("trt", float(cm.τ_symmetric), float(cm.τ_asymmetric))
            if key != self._coll_key:
                _lib.check(L.lbm_model_set_trt(self._h, key[1], key[2]), "lbm_model_set_trt")
        elif isinstance(cm, MRT):
            key = ("mrt",) + tuple(float(t) for t in cm.τs)
            if key != self._coll_key:
                t = np.ascontiguousarray(cm.τs, dtype=np.float64)
                _lib.check(L.lbm_model_set_mrt(self._h, _lib.dptr(t), t.size), "lbm_model_set_mrt")
        elif isinstance(cm, IterativeInitializationCollisionModel):
            # keyed by the model object, which the handle keeps alive (its
            # velocity is a private copy, fixed at construction)
            key = ("mei", id(cm))
            self._coll_obj = cm
            if key != self._coll_key:
                nx, ny, _ = cm.u.shape
                _lib.check(L.lbm_model_set_mei(self._h, cm.τ, _lib.dptr(cm.u), nx, ny),
                           "lbm_model_set_mei")
        else:
            raise TypeError(f"unsupported collision model {cm!r}")
        self._coll_key = key

    def set_force(self, force: Any, nx: int, ny: int, time: float) -> None:
        L = _lib.lib()
        if force is None:
            if self._force_key != ("none",):
                _lib.check(L.lbm_model_set_force(self._h, _lib.FORCE_NONE, None, 0, 0),
                           "lbm_model_set_force")
            self._force_key = ("none",)
            return
        if callable(force):
            # the reference's per-node closure (x_idx, y_idx, t) -> F, 1-based
            field = np.empty((nx, ny, 2), order="F")
            for x in range(nx):
                for y in range(ny):
                    F = force(x + 1, y + 1, time)
                    field[x, y, 0], field[x, y, 1] = np.ravel(F)[:2]
            F = field
        else:
            F = np.asarray(force, dtype=np.float64)
        if F.size == 2:
            v = np.ascontiguousarray(F.reshape(2))
            key = ("const", float(v[0]), float(v[1]))
            if key != self._force_key:
                _lib.check(L.lbm_model_set_force(self._h, _lib.FORCE_CONST, _lib.dptr(v), 0, 0),
                           "lbm_model_set_force")
            self._force_key = key
            return
        if F.shape != (nx, ny, 2):
            raise ValueError(f"force field shape {F.shape} does not match ({nx}, {ny}, 2)")
        Ff = np.asfortranarray(F)
        _lib.check(L.lbm_model_set_force(self._h, _lib.FORCE_FIELD, _lib.dptr(Ff), nx, ny),
                   "lbm_model_set_force")
        self._force_key = ("field", id(force), time)

    def set_bcs(self, bcs: Sequence[BoundaryCondition]) -> None:
        key = tuple(bcs)
        if key == self._bcs_key:
            return
        n = len(bcs)
        arr = (_lib.LbmBC * max(1, n))()
        for k, bc in enumerate(bcs):
            arr[k] = bc.to_abi()
        _lib.check(_lib.lib().lbm_model_set_bcs(self._h, arr, n), "lbm_model_set_bcs")
        self._bcs_key = key

    def set_launch(self, launch) -> None:
        """(threads per block along x, rows per block, nodes per thread,
        non-temporal) of the fused step; None = automatic.  A 3-tuple omits
        the rows (one row per block)."""
        if launch is not None and len(launch) == 3:
            launch = (launch[0], 1, launch[1], launch[2])
        key = tuple(launch) if launch is not None else (0, 0, 0, -1)
        if key != self._launch_key:
            _lib.check(_lib.lib().lbm_model_set_launch(self._h, *[int(k) for k in key]),
                       "lbm_model_set_launch")
        self._launch_key = key

    def configure(self, cm, f: HipLattice, bcs=None, time: float = 0.0) -> "NativeModel":
        if cm is not None:
            self.set_collision(cm)
            self.set_force(getattr(cm, "force", None), f.nx, f.ny, time)
        if bcs is not None:
            self.set_bcs(list(bcs))
        return self

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.lib().lbm_model_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_models: dict = {}


def native_model(q: Quadrature) -> NativeModel:
    m = _models.get(q.id)
    if m is None:
        m = _models[q.id] = NativeModel(q)
    return m


def _check_lattice(q: Quadrature, *fs: HipLattice) -> None:
    for f in fs:
        if not isinstance(f, HipLattice):
            raise TypeError(f"expected a HipLattice, got {type(f).__name__}")
        if f.quadrature != q:
            raise ValueError(f"lattice holds {f.quadrature!r} populations, not {q!r}")


def collide_(cm: CollisionModel, q: Quadrature, f_in: HipLattice = None, f_out: HipLattice = None,
             *, time: float = 0.0, f_old: HipLattice = None, f_new: HipLattice = None) -> None:
    """collide!(cm, q, f_in, f_out; time) / collide!(cm, q; time, f_new, f_old)"""
    if f_in is None:
        f_in, f_out = f_old, f_new
    _check_lattice(q, f_in, f_out)
    m = native_model(q).configure(cm, f_in, None, time)
    _lib.check(_lib.lib().lbm_collide(m.handle, f_in.handle, f_out.handle), "collide!")


def stream_(q: Quadrature, f: HipLattice = None, f_new: HipLattice = None, *,
            f_old: HipLattice = None) -> None:
    """stream!(q, f, f_new) / stream!(q; f_new, f_old): periodic pull (stream.jl:18-30)"""
    if f is None:
        f = f_old
    _check_lattice(q, f, f_new)
    m = native_model(q)
    _lib.check(_lib.lib().lbm_stream(m.handle, f.handle, f_new.handle), "stream!")


def apply_(bcs: Sequence[BoundaryCondition], q: Quadrature, f_new: HipLattice, f_old: HipLattice,
           *, time: float = 0.0) -> None:
    """apply!(bcs, q, f_new, f_old; time): ordered list, in place on f_new"""
    if isinstance(bcs, BoundaryCondition):
        bcs = [bcs]
    _check_lattice(q, f_new, f_old)
    m = native_model(q)
    m.set_bcs(list(bcs))
    _lib.check(_lib.lib().lbm_apply_bcs(m.handle, f_new.handle, f_old.handle), "apply!")


def stream_bc_(q: Quadrature, bcs: Sequence[BoundaryCondition], f: HipLattice,
               f_new: HipLattice) -> None:
    """stream! followed by apply!, fused into one pass."""
    _check_lattice(q, f, f_new)
    m = native_model(q)
    m.set_bcs(list(bcs))
    _lib.check(_lib.lib().lbm_stream_bc(m.handle, f.handle, f_new.handle), "stream!+apply!")


def synchronize() -> None:
    _lib.check(_lib.lib().lbm_synchronize(), "lbm_synchronize")


CLOCK_PROBES = 1024  # LBM_CLOCK_PROBES (include/lbm_hip.h)


def clock_sample():
    """Shader-clock and constant-rate counters sampled on the library's
    stream after the work issued so far, one row per probe block: (rows of
    (location key, shader cycles, constant ticks), the constant counter's
    rate in MHz).  The key names the compute unit the block ran on."""
    import ctypes

    import numpy as np

    out = np.zeros(3 * CLOCK_PROBES, dtype=np.uint64)
    mhz = ctypes.c_double()
    _lib.check(_lib.lib().lbm_clock_sample(out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                           ctypes.byref(mhz)), "lbm_clock_sample")
    return out.reshape(CLOCK_PROBES, 3), mhz.value


def average_clock_mhz(before, after):
    """Average shader clock (MHz) over the device work between two
    ``clock_sample`` results.  The shader-clock counters of different compute
    units are not aligned, so the two samples are paired per location key:
    each compute unit seen in both gives cycles / ticks * tick rate, and the
    result is the median over those units (None if none pairs)."""
    import numpy as np

    (a, tick), (b, _) = before, after

    def first(rows):
        seen = {}
        for key, cyc, ticks in rows.tolist():
            if key not in seen or ticks < seen[key][1]:
                seen[key] = (cyc, ticks)
        return seen

    sa, sb = first(a), first(b)
    rates = []
    for key in sa.keys() & sb.keys():
        dt = sb[key][1] - sa[key][1]
        dc = sb[key][0] - sa[key][0]
        if dt > 0 and dc > 0:
            rates.append(dc / dt * tick)
    return float(np.median(rates)) if rates else None
