"""On-device processing: TrackHydrodynamicErrors and the stopping criteria.

The reference's processing methods walk the populations node by node on the
host (processing_methods/track_hydrodynamic_errors.jl:52-242,
processing_methods/stopping_criteria/stopping_criteria.jl:17-115).  Here the
populations never leave the GPU: a row of error norms is one pass of a HIP
kernel over the lattice (``lbm_track_errors``) and the stopping criteria read
their sums from another (``lbm_velocity_stats``), both with FP64 sums in a
fixed order.  Only 11 (or 5) numbers come back to the host.

The analytic fields a row compares against live on the device as eight planes
(:class:`ExpectedFields`).  A call scales them by per-call affine factors, so
a problem whose fields change in time only by a factor (a decaying vortex)
uploads them once.
"""
from __future__ import annotations

import ctypes
import math
import os
from typing import List, NamedTuple, Optional, Sequence

import numpy as np

from . import _lib
from .lattice import HipLattice
from .model import ProcessingMethod
from .quadratures import Quadrature

# order of the expected planes
FIELDS = ("rho", "ux", "uy", "p", "sigma_xx", "sigma_xy", "sigma_yx", "sigma_yy")
IDENTITY_AFFINE = (0.0, 1.0) * 8


class TrackRow(NamedTuple):
    """One entry of TrackHydrodynamicErrors' ``df``
    (track_hydrodynamic_errors.jl:19-46, 205-221)."""
    timestep: int
    error_rho: float
    error_u: float
    error_p: float
    error_sigma_xx: float
    error_sigma_xy: float
    error_sigma_yy: float
    error_sigma_yx: float
    mass: float
    momentum: float
    energy: float


class ExpectedFields:
    """Eight FP64 planes (rho, u_x, u_y, p, sigma_xx, sigma_xy, sigma_yx,
    sigma_yy) on the device, each (nx, ny) column-major."""

    def __init__(self, nx: int, ny: int, planes: Optional[np.ndarray] = None):
        self.nx, self.ny = int(nx), int(ny)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().lbm_expected_create(ctypes.byref(h), self.nx, self.ny),
                   "lbm_expected_create")
        self._h = h
        if planes is not None:
            self.upload(planes)

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def upload(self, planes: np.ndarray) -> "ExpectedFields":
        a = np.asfortranarray(planes, dtype=np.float64)
        if a.shape != (self.nx, self.ny, 8):
            raise ValueError(f"expected planes of shape {(self.nx, self.ny, 8)}, got {a.shape}")
        _lib.check(_lib.lib().lbm_expected_upload(self._h, _lib.dptr(a)), "lbm_expected_upload")
        return self

    def free(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().lbm_expected_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class VelocityTracker:
    """The previous velocity field of VelocityConvergenceStoppingCriteria,
    kept on the device (stopping_criteria.jl:57-69); zero when created."""

    def __init__(self, nx: int, ny: int):
        self.nx, self.ny = int(nx), int(ny)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().lbm_velocity_tracker_create(ctypes.byref(h), self.nx, self.ny),
                   "lbm_velocity_tracker_create")
        self._h = h

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def free(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().lbm_velocity_tracker_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def track_errors(f: HipLattice, expected: ExpectedFields, tau: float, u_max: float,
                 delta: float = 1.0, affine: Sequence[float] = IDENTITY_AFFINE) -> np.ndarray:
    """The 11 numbers of one row: error_rho, error_u, error_p, error_sigma_xx,
    error_sigma_xy, error_sigma_yy, error_sigma_yx, mass, momentum, energy
    (track_hydrodynamic_errors.jl:205-221) and the number of nodes holding a
    non-finite population.  ``tau`` is the stress relaxation time s*nu."""
    if len(affine) != 16:
        raise ValueError("affine needs 16 numbers: (offset, factor) per expected plane")
    prm = _lib.LbmTrackParams()
    prm.tau = float(tau)
    prm.u_max = float(u_max)
    prm.delta = float(delta)
    for k, v in enumerate(affine):
        prm.affine[k] = float(v)
    out = np.zeros(_TRACK_NOUT)
    _lib.check(_lib.lib().lbm_track_errors(f.handle, expected.handle, ctypes.byref(prm),
                                           _lib.dptr(out)), "lbm_track_errors")
    return out


_TRACK_NOUT = 11


def velocity_stats(f: HipLattice, tracker: Optional[VelocityTracker] = None) -> np.ndarray:
    """mean u_x, mean u_y, sum |u - u_old|^2, sum |u_old|^2 (the tracker then
    holds u) and the number of non-finite nodes."""
    out = np.zeros(5)
    _lib.check(_lib.lib().lbm_velocity_stats(f.handle, None if tracker is None else tracker.handle,
                                             _lib.dptr(out)), "lbm_velocity_stats")
    return out


# --------------------------------------------------------------------------
# stopping criteria (processing_methods/stopping_criteria/)
# --------------------------------------------------------------------------
class StopCriteria:
    def should_stop_(self, q: Quadrature, f: HipLattice) -> bool:
        """should_stop!(::StopCriteria, q, f_in) = false (stopping_criteria.jl:16)"""
        return False


class NoStoppingCriteria(StopCriteria):
    pass


class MeanVelocityStoppingCriteria(StopCriteria):
    """Stops when the mean u_x changes by a relative amount below
    ``tolerance`` between two checks, or is NaN (stopping_criteria.jl:3-55)."""

    def __init__(self, old_mean_velocity: float, tolerance: float, problem=None):
        self.old_mean_velocity = float(old_mean_velocity)
        self.tolerance = float(tolerance)
        self.problem = problem

    def should_stop_(self, q: Quadrature, f: HipLattice) -> bool:
        u_mean = np.float64(velocity_stats(f)[0])
        with np.errstate(divide="ignore", invalid="ignore"):
            converged = abs(u_mean / np.float64(self.old_mean_velocity) - 1)
        if converged < self.tolerance:
            return True
        if math.isnan(u_mean):
            return True
        self.old_mean_velocity = float(u_mean)
        return False


class VelocityConvergenceStoppingCriteria(StopCriteria):
    """Stops when sqrt(sum |u - u_old|^2) / sum |u_old|^2 is below
    ``tolerance`` or NaN; the denominator is not square-rooted, as in the
    reference (stopping_criteria.jl:56-115)."""

    def __init__(self, tolerance: float, problem=None, nx: Optional[int] = None,
                 ny: Optional[int] = None):
        self.tolerance = float(tolerance)
        self.problem = problem
        nx = problem.NX if nx is None else nx
        ny = problem.NY if ny is None else ny
        self.old_velocity = VelocityTracker(nx, ny)

    def should_stop_(self, q: Quadrature, f: HipLattice) -> bool:
        s = velocity_stats(f, self.old_velocity)
        with np.errstate(divide="ignore", invalid="ignore"):
            converged = np.sqrt(np.float64(s[2])) / np.float64(s[3])
        if converged < self.tolerance:
            return True
        if math.isnan(converged):
            return True
        return False


class DensityConvergence(StopCriteria):
    """density_convergence.jl:1-17, including its loop over the single node
    (nx, ny): only that node's density is tracked."""

    def __init__(self, eps: float, nx: int, ny: int):
        self.eps = float(eps)
        self.rho_old = np.zeros((nx, ny))
        self.rho = np.zeros((nx, ny))

    def should_stop_(self, q: Quadrature, f: HipLattice) -> bool:
        nx, ny = self.rho_old.shape
        self.rho_old[nx - 1, ny - 1] = self.rho[nx - 1, ny - 1]
        pops = f.node(nx, ny)
        r = pops[0] + pops[1]
        for v in pops[2:]:
            r = r + v
        self.rho[nx - 1, ny - 1] = r
        d = float(np.linalg.norm(self.rho - self.rho_old))
        return d < self.eps or d > 100.0


def stop_criteria_for(problem) -> StopCriteria:
    """StopCriteria(problem) (stopping_criteria.jl:8-15), by the problem's
    ``kind``: Poiseuille 1e-12, Couette 1e-7, lid-driven cavity 1e-5, a static
    decaying shear flow 1e-8, anything else none."""
    kind = getattr(problem, "kind", None)
    tol = {"poiseuille": 1e-12, "couette": 1e-7, "lid_driven_cavity": 1e-5}.get(kind)
    if kind == "decaying_shear" and getattr(problem, "static", False):
        tol = 1e-8
    return NoStoppingCriteria() if tol is None else MeanVelocityStoppingCriteria(0.0, tol, problem)


# --------------------------------------------------------------------------
# TrackHydrodynamicErrors (processing_methods/track_hydrodynamic_errors.jl)
# --------------------------------------------------------------------------
class TrackHydrodynamicErrors(ProcessingMethod):
    """Collects a :class:`TrackRow` at the steps where the reference's
    ``next!`` does (track_hydrodynamic_errors.jl:52-66): every step if
    ``should_process``, else at step ``n_steps`` and when the stopping
    criteria (checked every 100 steps) fire.  Plotting is not done.

    ``problem`` provides ``range()`` (cell centres), ``delta_t()``,
    ``lattice_viscosity()``, ``u_max``, ``NX``, ``NY`` and either
    ``expected_planes(q)`` with ``expected_affine(q, time)`` (planes uploaded
    once) or ``expected_fields(q, time)`` (uploaded at every row); the planes
    are (NX, NY, 8) in the order of :data:`FIELDS`."""

    def __init__(self, problem, should_process: bool, n_steps: int,
                 stop_criteria: Optional[StopCriteria] = None):
        self.problem = problem
        self.should_process = bool(should_process)
        self.n_steps = int(n_steps)
        self.stop_criteria = stop_criteria_for(problem) if stop_criteria is None else stop_criteria
        self.df: List[TrackRow] = []
        self._expected: Optional[ExpectedFields] = None

    def needs_step(self, t: int) -> bool:
        """Whether ``next_`` at step t reads the populations; at other steps it
        returns False without side effects, so the steps in between may run
        fused."""
        return self.should_process or t % 100 == 0 or t == self.n_steps

    def _cell_area(self) -> float:
        # Δ_ (track_hydrodynamic_errors.jl:85-94)
        xr, yr = self.problem.range()
        nx, ny = len(xr), len(yr)
        dx = float(xr[-1] - xr[0]) / (nx - 1) if nx > 1 else 0.0
        dy = float(yr[-1] - yr[0]) / (ny - 1) if ny > 1 else 0.0
        if nx == 1:
            return 1.0 if ny == 1 else dy
        if ny == 1:
            return dx
        return dy * dx

    def row(self, q: Quadrature, f: HipLattice, t: int) -> TrackRow:
        p = self.problem
        time = t * p.delta_t()
        if hasattr(p, "expected_planes"):
            if self._expected is None:
                self._expected = ExpectedFields(f.nx, f.ny, p.expected_planes(q))
            affine = p.expected_affine(q, time)
        else:
            if self._expected is None:
                self._expected = ExpectedFields(f.nx, f.ny)
            self._expected.upload(p.expected_fields(q, time))
            affine = IDENTITY_AFFINE
        tau = q.speed_of_sound_squared * p.lattice_viscosity()
        out = track_errors(f, self._expected, tau, p.u_max, self._cell_area(), affine)
        return TrackRow(int(t), *[float(v) for v in out[:10]])

    def next_(self, q: Quadrature, f: HipLattice, t: int) -> bool:
        should_stop = False
        if t % 100 == 0 and self.stop_criteria.should_stop_(q, f):
            should_stop = True
        if not should_stop and t != self.n_steps and not self.should_process:
            return False
        self.df.append(self.row(q, f, t))
        return should_stop


# --------------------------------------------------------------------------
# TakeSnapshots (processing_methods/take_snapshots.jl)
# --------------------------------------------------------------------------
class TakeSnapshots(ProcessingMethod):
    """TakeSnapshots(problem, every_t) (take_snapshots.jl:3-29): at every
    ``every_t``-th step (an int) or at the listed steps (a list) the
    populations are copied (``snapshot = copy(f_in)``).  The copy goes through
    one of ``device_slots`` staging buffers in HBM to pinned host memory on a
    second stream (lbm_snapshot_take), so the step does not wait for the
    host; with ``directory`` each snapshot is also written there as
    snapshot_<t>.npy.  :meth:`arrays` gives the (nx, ny, Q) Float64 arrays.
    Plotting is not done."""

    def __init__(self, problem, every_t, device_slots: int = 2, directory: str = None):
        self.problem = problem
        self.every_t = every_t
        self.device_slots = int(device_slots)
        self.directory = directory
        self.timesteps: List[int] = []
        self._h = None
        self._shape = None

    def needs_step(self, t: int) -> bool:
        if isinstance(self.every_t, int):
            return t % self.every_t == 0
        return t in self.every_t

    def _handle(self, q: Quadrature, f: HipLattice):
        import ctypes
        if self._h is None:
            h = ctypes.c_void_p()
            d = None if self.directory is None else os.fsencode(self.directory)
            _lib.check(_lib.lib().lbm_snapshots_create(ctypes.byref(h), q.id, f.nx, f.ny,
                                                       self.device_slots, d),
                       "lbm_snapshots_create")
            self._h = h
            self._shape = (f.nx, f.ny, q.Q)
        return self._h

    def next_(self, q: Quadrature, f: HipLattice, t: int) -> bool:
        if not self.needs_step(t):
            return False
        _lib.check(_lib.lib().lbm_snapshot_take(self._handle(q, f), f.handle, int(t)),
                   "lbm_snapshot_take")
        self.timesteps.append(int(t))
        return False

    def wait(self) -> int:
        import ctypes
        if self._h is None:
            return 0
        n = ctypes.c_int()
        _lib.check(_lib.lib().lbm_snapshots_wait(self._h, ctypes.byref(n)), "lbm_snapshots_wait")
        return n.value

    def arrays(self) -> List[np.ndarray]:
        """the snapshots as (nx, ny, Q) Float64 host arrays, in the order taken"""
        import ctypes
        out = []
        for k in range(self.wait()):
            a = np.empty(self._shape, order="F")
            t = ctypes.c_int64()
            _lib.check(_lib.lib().lbm_snapshot_get(self._h, k, ctypes.byref(t), _lib.dptr(a)),
                       "lbm_snapshot_get")
            out.append(a)
        return out

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            try:
                _lib.lib().lbm_snapshots_destroy(self._h)
            except Exception:
                pass
            self._h = None
