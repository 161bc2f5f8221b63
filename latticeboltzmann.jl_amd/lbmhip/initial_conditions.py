"""Initial populations (src/initial_conditions.jl:7-22, initial_conditions/).

``initialize(strategy, q, problem)`` runs the reference's strategies on the
device: AnalyticalEquilibrium and AnalyticalEquilibriumAndOffEquilibrium per
node from the problem's lattice fields (lbm_initialize_analytical, the
Hermite sum of velocity_distribution_function/hermite.jl:10-33 in the
reference's order), and the iterative initialization of Mei et al.
``equilibrium_populations(q, ρ, ux, uy)`` is a host convenience for tests and
benchmarks: the closed-form equilibrium polynomial, equal to the Hermite sum
to rounding.
"""
from __future__ import annotations

import numpy as np

from .quadratures import Quadrature


def equilibrium_populations(q: Quadrature, rho, ux, uy) -> np.ndarray:
    rho = np.asarray(rho, dtype=np.float64)
    ux = np.asarray(ux, dtype=np.float64)
    uy = np.asarray(uy, dtype=np.float64)
    s = q.speed_of_sound_squared
    usq = ux * ux + uy * uy
    fo = {4: 1, 5: 1, 9: 2, 13: 2, 17: 3, 21: 3, 37: 4}[q.id]
    out = np.empty(rho.shape + (q.Q,), order="F")
    for i in range(q.Q):
        ud = q.abscissae[0, i] * ux + q.abscissae[1, i] * uy
        a1 = s * ud
        acc = 1.0 + a1
        if fo >= 2:
            acc = acc + 0.5 * ((s * s) * (ud * ud) - s * usq)
        if fo >= 3:
            acc = acc + (1.0 / 6.0) * ((s * ud) * ((s * s) * (ud * ud) - 3.0 * s * usq))
        if fo >= 4:
            acc = acc + (1.0 / 24.0) * ((s ** 4) * ud ** 4 - 6.0 * s ** 3 * usq * ud * ud
                                        + 3.0 * s * s * usq * usq)
        out[..., i] = (rho * q.weights[i]) * acc
    return out


def cell_centres(n: int, length: float = 1.0) -> np.ndarray:
    """range(Δx/2, stop = L - Δx/2, length = n) -- problems.jl:18-26"""
    dx = length / n
    return np.linspace(dx / 2, length - dx / 2, n)


class IterativeInitializationMeiEtAl:
    """IterativeInitializationMeiEtAl(τ, ϵ) (initial_conditions/mei_et_al.jl:
    1-10); IterativeInitialization() = (1.0, 1e-7)."""

    def __init__(self, τ: float = 1.0, ϵ: float = 1e-7):
        self.τ = float(τ)
        self.ϵ = float(ϵ)


def initialize_mei(strategy: IterativeInitializationMeiEtAl, q: Quadrature, problem,
                   max_steps: int = 10000):
    """initialize(::IterativeInitializationMeiEtAl, q, problem) (mei_et_al.jl:
    13-41) on the device: f = w everywhere, then the collide-only model of the
    prescribed velocity, stream! and the problem's boundary conditions, until
    the density at node (nx, ny) changes by less than ϵ (or more than 100)
    between steps, at most max_steps (the reference's 1:10000).  ``problem``
    gives ``NX``, ``NY``, ``lattice_velocity_field()`` and optionally
    ``boundary_conditions()``.  Returns (f_stream as a HipLattice, steps)."""
    import ctypes

    from . import _lib
    from .collision_models import IterativeInitializationCollisionModel
    from .lattice import HipLattice
    from .operators import native_model

    cm = IterativeInitializationCollisionModel.for_problem(q, strategy.τ, problem)
    bcs = list(problem.boundary_conditions()) if hasattr(problem, "boundary_conditions") else []
    fs = HipLattice(q, problem.NX, problem.NY)
    fc = HipLattice(q, problem.NX, problem.NY)
    m = native_model(q).configure(cm, fs, bcs)
    steps = ctypes.c_int64()
    _lib.check(_lib.lib().lbm_mei_initialize(m.handle, fs.handle, fc.handle, strategy.ϵ,
                                             int(max_steps), ctypes.byref(steps)),
               "lbm_mei_initialize")
    return fs, int(steps.value)


class AnalyticalEquilibrium:
    """AnalyticalEquilibrium() (initial_conditions/analytical_equilibrium.jl):
    f0 = feq(rho(x), u(x), T(x)) of the problem's analytic fields, by the
    Hermite sum of hermite_based_equilibrium! (hermite.jl:10-33)."""


class AnalyticalEquilibriumAndOffEquilibrium:
    """AnalyticalEquilibriumAndOffEquilibrium()
    (initial_conditions/analytical_offequilibrium.jl:9): the analytic
    equilibrium plus the first-order non-equilibrium part built from the
    problem's velocity gradient."""


# lbm_initialize_analytical strategies (include/lbm_hip.h)
INIT_EQUILIBRIUM, INIT_OFFEQ_TGV, INIT_OFFEQ = 0, 1, 2


def initialize_analytical(q: Quadrature, rho, ux, uy, T=None, grad=None,
                          strategy: int = INIT_EQUILIBRIUM, tau: float = 0.0,
                          factor: float = 1.0, storage: str = "f64"):
    """The analytic initial conditions computed on the device from (nx, ny)
    planes of lattice density, velocity and (optionally) temperature, and for
    the off-equilibrium strategies the (nx, ny, 2, 2) velocity gradient
    G[..., a, b] as the reference passes it; tau = s nu.  Returns a
    HipLattice in ``storage``."""
    from . import _lib
    from .lattice import HipLattice

    rho = np.asfortranarray(rho, dtype=np.float64)
    nx, ny = rho.shape
    ux = np.asfortranarray(np.broadcast_to(ux, (nx, ny)), dtype=np.float64)
    uy = np.asfortranarray(np.broadcast_to(uy, (nx, ny)), dtype=np.float64)
    Tp = None if T is None else np.asfortranarray(np.broadcast_to(T, (nx, ny)), dtype=np.float64)
    Gp = None
    if grad is not None:
        G = np.asarray(grad, dtype=np.float64)
        if G.shape != (nx, ny, 2, 2):
            raise ValueError(f"velocity gradient shape {G.shape} is not ({nx}, {ny}, 2, 2)")
        Gp = np.asfortranarray(G.reshape(nx, ny, 4, order="F"))
    f = HipLattice(q, nx, ny, storage=storage)
    _lib.check(_lib.lib().lbm_initialize_analytical(
        f.handle, int(strategy), _lib.dptr(rho), _lib.dptr(ux), _lib.dptr(uy),
        None if Tp is None else _lib.dptr(Tp), None if Gp is None else _lib.dptr(Gp),
        float(tau), float(factor)), "lbm_initialize_analytical")
    return f


def initialize(strategy, q: Quadrature, problem, storage: str = "f64"):
    """initialize(strategy, q, problem) (initial_conditions.jl:7-22,
    mei_et_al.jl:13-41) returning the populations on the device.  The problem
    supplies its analytic lattice fields (``lattice_fields(q)``: density,
    velocity and temperature planes; ``offequilibrium_inputs()``: the
    strategy variant, the velocity gradient and the area factor for the
    off-equilibrium start); the per-node initial conditions are evaluated on
    the device (lbm_initialize_analytical); the iterative initialization runs
    on the device."""
    if isinstance(strategy, IterativeInitializationMeiEtAl):
        if storage != "f64":
            raise ValueError("the iterative initialization runs on FP64 storage only")
        return initialize_mei(strategy, q, problem)[0]
    if isinstance(strategy, AnalyticalEquilibrium):
        rho, ux, uy, T = problem.lattice_fields(q)
        return initialize_analytical(q, rho, ux, uy, T, storage=storage)
    if isinstance(strategy, AnalyticalEquilibriumAndOffEquilibrium):
        rho, ux, uy, T = problem.lattice_fields(q)
        mode, grad, factor = problem.offequilibrium_inputs()
        tau = q.speed_of_sound_squared * problem.lattice_viscosity()
        return initialize_analytical(q, rho, ux, uy, T, grad, mode, tau, factor, storage=storage)
    raise TypeError(f"unsupported initialization strategy {strategy!r}")
