"""Macroscopic fields and FP64 global sums computed on the GPU.

``moments(f, τ)`` returns per-node ρ, u and the deviatoric stress σ as the
reference computes them per node in TrackHydrodynamicErrors
(track_hydrodynamic_errors.jl:139-177): ``density`` / ``velocity!``
(moments.jl:1-19) and ``deviatoric_tensor(q, τ, f, ρ, u)`` (moments.jl:81-96).
``conserved_sums(f)`` gives the FP64 mass / momentum / energy sums and the
number of nodes holding a non-finite population (the NaN check of
stopping_criteria.jl:47-50).
"""
from __future__ import annotations

from typing import NamedTuple

import numpy as np

from . import _lib
from .lattice import HipLattice


class Moments(NamedTuple):
    rho: np.ndarray
    ux: np.ndarray
    uy: np.ndarray
    sigma_xx: np.ndarray
    sigma_xy: np.ndarray
    sigma_yy: np.ndarray


class Sums(NamedTuple):
    mass: float
    momentum_x: float
    momentum_y: float
    energy: float
    nonfinite_nodes: int


def moments(f: HipLattice, τ: float = 1.0, stress: bool = True) -> Moments:
    outs = [np.empty((f.nx, f.ny), order="F") for _ in range(6)]
    ptrs = [_lib.dptr(o) for o in outs[:3]] + ([_lib.dptr(o) for o in outs[3:]] if stress
                                               else [None, None, None])
    _lib.check(_lib.lib().lbm_moments(f.handle, float(τ), *ptrs), "lbm_moments")
    if not stress:
        outs[3:] = [None, None, None]
    return Moments(*outs)


def density(f: HipLattice) -> np.ndarray:
    return moments(f, stress=False).rho


def velocity(f: HipLattice):
    m = moments(f, stress=False)
    return m.ux, m.uy


def conserved_sums(f: HipLattice) -> Sums:
    out = np.zeros(5)
    _lib.check(_lib.lib().lbm_reduce(f.handle, _lib.dptr(out)), "lbm_reduce")
    return Sums(float(out[0]), float(out[1]), float(out[2]), float(out[3]), int(out[4]))
