"""Boundary conditions (src/boundary_conditions.jl, boundary_conditions/*.jl).

``BounceBack(direction, xs, ys)`` -- halfway bounce-back (bounce_back.jl:1-72);
``MovingWall(direction, xs, ys, u, ρ=1, T=1)`` -- bounce-back plus the wall
momentum term 2 w_i s ρ_w u_w·c_i (moving_wall.jl:1-38; North only, and,
as in the reference, it acts on the whole lattice regardless of xs, ys).
Ranges are Julia UnitRanges: pass ``range(1, NX + 1)`` or an inclusive
``(first, last)`` pair, 1-based.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence, Tuple

from . import _lib


class Direction:
    code: int = -1

    def __repr__(self):
        return type(self).__name__ + "()"

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return self.code


class North(Direction):
    code = _lib.NORTH


class East(Direction):
    code = _lib.EAST


class South(Direction):
    code = _lib.SOUTH


class West(Direction):
    code = _lib.WEST


def _unit_range(r) -> Tuple[int, int]:
    if isinstance(r, range):
        if r.step != 1:
            raise ValueError("boundary ranges must be unit ranges")
        return (r.start, r.stop - 1)
    a, b = r
    return (int(a), int(b))


class BoundaryCondition:
    def to_abi(self) -> _lib.LbmBC:
        raise NotImplementedError


@dataclass(frozen=True)
class BounceBack(BoundaryCondition):
    direction: Direction
    xs: Tuple[int, int]
    ys: Tuple[int, int]

    def __init__(self, direction, xs, ys):
        object.__setattr__(self, "direction", direction)
        object.__setattr__(self, "xs", _unit_range(xs))
        object.__setattr__(self, "ys", _unit_range(ys))

    def to_abi(self):
        return _lib.LbmBC(_lib.BOUNCE_BACK, self.direction.code, self.xs[0], self.xs[1],
                          self.ys[0], self.ys[1], 0.0, 0.0, 1.0)


@dataclass(frozen=True)
class MovingWall(BoundaryCondition):
    direction: Direction
    xs: Tuple[int, int]
    ys: Tuple[int, int]
    u: Tuple[float, float]
    ρ: float = 1.0
    T: float = 1.0

    def __init__(self, direction, xs, ys, u: Sequence[float], ρ: float = 1.0, T: float = 1.0):
        if not isinstance(direction, North):
            raise NotImplementedError("MovingWall is defined for North only (moving_wall.jl:17)")
        object.__setattr__(self, "direction", direction)
        object.__setattr__(self, "xs", _unit_range(xs))
        object.__setattr__(self, "ys", _unit_range(ys))
        object.__setattr__(self, "u", (float(u[0]), float(u[1])))
        object.__setattr__(self, "ρ", float(ρ))
        object.__setattr__(self, "T", float(T))

    def to_abi(self):
        return _lib.LbmBC(_lib.MOVING_WALL, self.direction.code, self.xs[0], self.xs[1],
                          self.ys[0], self.ys[1], self.u[0], self.u[1], self.ρ)
