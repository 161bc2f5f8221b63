"""Collision models SRT, TRT, MRT (src/collision_models.jl, collision_models/*.jl).

Constructors mirror the reference, including its argument-order quirk:
``TRT(τ_a, τ_s)`` with two arguments swaps them (trt.jl:6), while the
three-field form ``TRT.fields(τ_s, τ_a, force)`` is the struct order
(trt.jl:1-5).

Forces.  The reference stores a Julia closure ``(x_idx, y_idx, t) -> F`` and
calls it per node (srt.jl:52-54).  A closure cannot cross the C ABI, so the host
evaluates it into a lattice force field; a constant 2-vector or a prebuilt
``(nx, ny, 2)`` array can be given directly instead.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Optional, Sequence

import numpy as np

from .quadratures import Quadrature, order


class CollisionModel:
    force: Any = None


@dataclass
class SRT(CollisionModel):
    """SRT{Force}(τ, force) -- srt.jl:1-5"""
    τ: float
    force: Any = None

    @property
    def tau(self) -> float:
        return self.τ


@dataclass
class TRT(CollisionModel):
    """TRT{Force,T}(τ_symmetric, τ_asymmetric, force) -- trt.jl:1-5.

    ``TRT(a, b)`` follows trt.jl:6: ``TRT(τ_a, τ_s) = TRT(τ_s, τ_a, nothing)``.
    """
    τ_symmetric: float
    τ_asymmetric: float
    force: Any = None

    def __init__(self, *args, force=None):
        if len(args) == 2:
            τ_a, τ_s = args
            self.τ_symmetric, self.τ_asymmetric, self.force = float(τ_s), float(τ_a), force
        elif len(args) == 3:
            self.τ_symmetric, self.τ_asymmetric, self.force = float(args[0]), float(args[1]), args[2]
        else:
            raise TypeError("TRT(τ_a, τ_s) or TRT(τ_s, τ_a, force)")

    @classmethod
    def fields(cls, τ_symmetric, τ_asymmetric, force=None) -> "TRT":
        return cls(τ_symmetric, τ_asymmetric, force)


def trt_magic(τ: float, Λ: float = 0.25, force=None) -> TRT:
    """Magic-parameter factory: τ_s = τ, τ_a = 1/2 + Λ/(τ - 1/2) (trt.jl:8-21)."""
    τ_ = 0.5 + Λ / (τ - 0.5)
    return TRT.fields(τ, τ_, force)


@dataclass
class TRT_Λ:
    """TRT_Λ(Λ) -- trt.jl:35-40"""
    Λ: float


@dataclass
class MRT(CollisionModel):
    """Regularised MRT (mrt.jl:1-34): ``MRT(q, τ)`` fills
    round(order/2) relaxation times; ``MRT(q, τs)`` takes them explicitly;
    ``MRT(q, τ_s, τ_a)`` alternates them (mrt.jl:24-27)."""
    τs: Sequence[float] = field(default_factory=list)
    force: Any = None

    def __init__(self, q: Quadrature, *taus, force=None):
        N = int(round(order(q) / 2.0))  # Julia round: ties to even (mrt.jl:20)
        if len(taus) == 1 and np.ndim(taus[0]) == 0:
            ts = [float(taus[0])] * N
        elif len(taus) == 1:
            ts = [float(t) for t in taus[0]]
        elif len(taus) == 2:
            ts = ([float(taus[0]), float(taus[1])] * N)[:N]
        else:
            raise TypeError("MRT(q, τ), MRT(q, τs) or MRT(q, τ_s, τ_a)")
        self.τs = ts
        self.force = force


class IterativeInitializationCollisionModel(CollisionModel):
    """IterativeInitializationCollisionModel(q, τ, problem)
    (collision_models/iterative_initialization.jl:1-40): relaxes towards
    w_i ρ + N_i, where N_i, the nonlinear part of the equilibrium of the
    prescribed lattice velocity, is formed once (on the host, in the
    reference's order) and kept on the device.  ``u`` is that velocity,
    (nx, ny, 2); ``IterativeInitializationCollisionModel.for_problem(q, τ,
    problem)`` takes it from ``problem.lattice_velocity_field()``."""

    def __init__(self, τ: float, u):
        self.τ = float(τ)
        self.u = np.array(u, dtype=np.float64, order="F", copy=True)
        if self.u.ndim != 3 or self.u.shape[2] != 2:
            raise ValueError(f"lattice velocity must be (nx, ny, 2), got {self.u.shape}")
        self.force = None

    @classmethod
    def for_problem(cls, q: Quadrature, τ: float, problem):
        return cls(τ, problem.lattice_velocity_field())


def collision_model_from_viscosity(kind, q: Quadrature, ν: float, force=None, Λ: float = 0.25):
    """CollisionModel(SRT|TRT|MRT, q, problem) with τ = s ν + 1/2
    (collision_models.jl:11-18, srt.jl:7-16, trt.jl:8-21, mrt.jl:36-47)."""
    τ = q.speed_of_sound_squared * ν + 0.5
    if kind is SRT:
        return SRT(τ, force)
    if kind is TRT or isinstance(kind, TRT_Λ):
        lam = kind.Λ if isinstance(kind, TRT_Λ) else Λ
        return trt_magic(τ, lam, force)
    if kind is MRT:
        return MRT(q, [τ] * order(q), force=force)
    raise TypeError(f"unknown collision model {kind!r}")
