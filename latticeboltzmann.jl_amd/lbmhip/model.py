"""LatticeBoltzmannModel and simulate (src/lattice_boltzmann_model.jl).

The model owns two device population arrays by role, as in the reference
(lattice_boltzmann_model.jl:1-14): ``f_stream`` (post-stream, post-BC; the
canonical state) and ``f_collision`` (post-collision).  ``simulate`` runs the
reference's step loop (lattice_boltzmann_model.jl:60-77)::

    for t in time
        collide!(model; time = t*Δt)   # f_stream -> f_collision
        stream!(model)                 # f_collision -> f_stream
        apply_boundary_conditions!(model; time = t*Δt)
        next!(model, t + 1) && return model
    end
    next!(model, last(time) + 1)

Without a processing method, and with a force that does not depend on time,
the whole loop is one ``lbm_simulate`` call: one fused pull collide-stream
kernel per step, after which both buffers hold what the reference's hold.
"""
from __future__ import annotations

from typing import Callable, Optional, Sequence

import numpy as np

from . import _lib
from .boundary_conditions import BoundaryCondition
from .collision_models import CollisionModel
from .lattice import HipLattice
from .operators import apply_, collide_, native_model, stream_
from .quadratures import Quadrature


class ProcessingMethod:
    """Per-step hook: ``next_(q, f_stream, t) -> bool`` (True stops the run);
    processing_methods.jl:92.  A method that defines ``needs_step(t)`` promises
    that ``next_`` at the other steps returns False and changes nothing; the
    steps in between then run as fused launches."""

    def next_(self, q: Quadrature, f: HipLattice, t: int) -> bool:
        return False


class LatticeBoltzmannModel:
    def __init__(self, f_stream: HipLattice, f_collision: HipLattice, quadrature: Quadrature,
                 collision_model: CollisionModel,
                 boundary_conditions: Sequence[BoundaryCondition] = (),
                 processing_method: Optional[ProcessingMethod] = None, delta_t: float = 0.0):
        self.f_stream = f_stream
        self.f_collision = f_collision
        self.quadrature = quadrature
        self.collision_model = collision_model
        self.boundary_conditions = list(boundary_conditions)
        self.processing_method = processing_method
        self.delta_t = float(delta_t)
        # launch shape of the fused step: None (automatic) or (threads per
        # block along x, rows per block, nodes per thread, non-temporal);
        # see autotune()
        self.launch = None

    @classmethod
    def from_populations(cls, q: Quadrature, f0: np.ndarray, collision_model: CollisionModel,
                         boundary_conditions: Sequence[BoundaryCondition] = (),
                         processing_method: Optional[ProcessingMethod] = None,
                         storage: str = "f64", rho0: float = 1.0,
                         delta_t: float = 0.0) -> "LatticeBoltzmannModel":
        """LatticeBoltzmannModel(problem, q; ...) given the initial populations
        (f_collision = copy(f_stream), lattice_boltzmann_model.jl:22-23)."""
        fs = HipLattice.from_numpy(q, f0, storage=storage, rho0=rho0)
        return cls(fs, fs.copy(), q, collision_model, boundary_conditions, processing_method,
                   delta_t)


def collide_model_(model: LatticeBoltzmannModel, time: float = 0.0) -> None:
    """collide!(model; time) -- lattice_boltzmann_model.jl:84-92"""
    collide_(model.collision_model, model.quadrature, f_new=model.f_collision,
             f_old=model.f_stream, time=time)


def stream_model_(model: LatticeBoltzmannModel) -> None:
    """stream!(model) -- lattice_boltzmann_model.jl:94-96"""
    stream_(model.quadrature, f_new=model.f_stream, f_old=model.f_collision)


def apply_boundary_conditions_(model: LatticeBoltzmannModel, time: float = 0.0) -> None:
    """apply_boundary_conditions!(model; time) -- lattice_boltzmann_model.jl:98-106"""
    if model.boundary_conditions:
        apply_(model.boundary_conditions, model.quadrature, model.f_stream, model.f_collision,
               time=time)


def next_(model: LatticeBoltzmannModel, t: int) -> bool:
    """next!(model, t) -- lattice_boltzmann_model.jl:108-110"""
    pm = model.processing_method
    if pm is None:
        return False
    return bool(pm.next_(model.quadrature, model.f_stream, t))


def simulate(model: LatticeBoltzmannModel, time) -> LatticeBoltzmannModel:
    """simulate(model, time) -- lattice_boltzmann_model.jl:60-77.  ``time`` is
    a range of integer time steps (Julia ``0:n`` is ``range(0, n + 1)``)."""
    steps = list(time) if not isinstance(time, range) else time
    if len(steps) == 0:
        next_(model, 0)
        return model
    force = getattr(model.collision_model, "force", None)
    pm = model.processing_method
    needs = getattr(pm, "needs_step", None) if pm is not None else None
    if not callable(force) and (pm is None or callable(needs)):
        # fused steps, broken only where the processing method reads the state
        m = native_model(model.quadrature).configure(model.collision_model, model.f_stream,
                                                     model.boundary_conditions,
                                                     float(steps[0]) * model.delta_t)
        m.set_launch(model.launch)
        n = len(steps)
        i = 0
        while i < n:
            j = i
            while pm is not None and j < n and not needs(int(steps[j]) + 1):
                j += 1
            if pm is None:
                j = n
            count = min(j, n - 1) - i + 1
            _lib.check(_lib.lib().lbm_simulate(m.handle, model.f_stream.handle,
                                               model.f_collision.handle, count), "simulate")
            if j < n and next_(model, int(steps[j]) + 1):
                return model
            i += count
        if pm is not None:
            next_(model, int(steps[-1]) + 1)
        return model
    m = native_model(model.quadrature)
    for t in steps:
        now = float(t) * model.delta_t
        m.configure(model.collision_model, model.f_stream, model.boundary_conditions, now)
        _lib.check(_lib.lib().lbm_collide(m.handle, model.f_stream.handle,
                                          model.f_collision.handle), "collide!")
        _lib.check(_lib.lib().lbm_stream_bc(m.handle, model.f_collision.handle,
                                            model.f_stream.handle), "stream!+apply!")
        if next_(model, int(t) + 1):
            return model
    next_(model, int(steps[-1]) + 1)
    return model


def simulate_timed(model: LatticeBoltzmannModel, nsteps: int):
    """As ``simulate(model, range(nsteps))`` without processing; returns
    (milliseconds spent in the fused kernels, number of fused launches),
    measured with HIP events on the library's stream."""
    import ctypes

    m = native_model(model.quadrature).configure(model.collision_model, model.f_stream,
                                                 model.boundary_conditions, 0.0)
    m.set_launch(model.launch)
    ms = ctypes.c_double()
    nl = ctypes.c_int64()
    _lib.check(_lib.lib().lbm_simulate_timed(m.handle, model.f_stream.handle,
                                             model.f_collision.handle, int(nsteps),
                                             ctypes.byref(ms), ctypes.byref(nl)), "simulate")
    return ms.value, nl.value


def autotune(model: LatticeBoltzmannModel, trial_steps: int = 20) -> dict:
    """Time every launch shape of the fused step on scratch copies of the
    model's lattice (the model state is not advanced), keep the fastest in
    ``model.launch`` and return it with its time per step."""
    import ctypes

    m = native_model(model.quadrature).configure(model.collision_model, model.f_stream,
                                                 model.boundary_conditions, 0.0)
    b, r, v, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    us = ctypes.c_double()
    _lib.check(_lib.lib().lbm_autotune(m.handle, model.f_stream.handle, int(trial_steps),
                                       ctypes.byref(b), ctypes.byref(r), ctypes.byref(v),
                                       ctypes.byref(n), ctypes.byref(us)), "lbm_autotune")
    model.launch = (b.value, r.value, v.value, n.value)
    m._launch_key = model.launch
    return {"block": b.value, "rows_per_block": r.value, "nodes_per_thread": v.value,
            "nontemporal": n.value, "us_per_step": us.value}


def simulate_launch_times(model: LatticeBoltzmannModel, nsteps: int):
    """As ``simulate(model, range(nsteps))`` without processing, with a HIP
    event before every fused launch: returns the per-launch milliseconds
    (nsteps - 1 values), for the spread of a kernel's duration over a run."""
    import ctypes

    import numpy as np

    m = native_model(model.quadrature).configure(model.collision_model, model.f_stream,
                                                 model.boundary_conditions, 0.0)
    m.set_launch(model.launch)
    out = np.zeros(max(int(nsteps) - 1, 1))
    nl = ctypes.c_int64()
    _lib.check(_lib.lib().lbm_simulate_launch_times(m.handle, model.f_stream.handle,
                                                    model.f_collision.handle, int(nsteps),
                                                    _lib.dptr(out), ctypes.byref(nl)), "simulate")
    return out[:nl.value]
