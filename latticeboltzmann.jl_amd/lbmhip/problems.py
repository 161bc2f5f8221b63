"""Analytic problems that feed the on-device processing.

Only the Taylor-Green vortex of problems/taylor_green_vortex.jl is mirrored:
its analytic fields (density = pressure, velocity, velocity gradient, the
deviatoric stress of problems.jl:35-50), the lattice scalings of problems.jl
(delta_x, delta_t, viscosity, the cell-centred range) and, for the processing
kernel, the split of every field into a fixed plane times a factor of time
(the vortex decays as exp(-(a^2 + b^2) nu t) in velocity and stress, and as
its square in pressure).
"""
from __future__ import annotations

import math

import numpy as np

from .initial_conditions import equilibrium_populations
from .quadratures import Quadrature


class TaylorGreenVortex:
    """TaylorGreenVortex(ν = 1/6, scale = 2, NX = 16 scale, NY = NX,
    domain_size = (2π, 2π); static = true, A = 1, B = -1, a = 1, b = 1)
    (taylor_green_vortex.jl:1-34)."""

    kind = "taylor_green_vortex"

    def __init__(self, ν: float = 1.0 / 6.0, scale: int = 2, NX: int = None, NY: int = None,
                 domain_size=(2 * math.pi, 2 * math.pi), static: bool = True, A: float = 1.0,
                 B: float = -1.0, a: float = 1.0, b: float = 1.0):
        self.NX = 16 * scale if NX is None else int(NX)
        self.NY = self.NX if NY is None else int(NY)
        self.rho_0 = 1.0
        self.u_max = 0.01 / scale
        self.ν = float(ν)
        self.domain_size = tuple(float(v) for v in domain_size)
        self.static = bool(static)
        self.A, self.B, self.a, self.b = float(A), float(B), float(a), float(b)

    # -- lattice scalings (problems.jl:79-107) --------------------------------
    def delta_x(self) -> float:
        if self.NX > self.NY:
            return self.domain_size[0] * (1 / self.NX)
        return self.domain_size[1] * (1 / self.NY)

    def delta_t(self) -> float:
        return self.delta_x() * self.u_max

    def viscosity(self) -> float:
        return self.ν * self.delta_x() ** 2 / self.delta_t()

    def lattice_viscosity(self) -> float:
        return self.ν

    def range(self):
        """Cell centres (problems.jl:18-26); y stops at domain_size[1] - Δy/2
        with the x extent, as in the reference."""
        dx = self.domain_size[0] / self.NX
        dy = self.domain_size[1] / self.NY
        x = np.linspace(dx / 2, self.domain_size[0] - dx / 2, self.NX)
        y = np.linspace(dy / 2, self.domain_size[0] - dy / 2, self.NY)
        return x, y

    def decay(self, time: float) -> float:
        if self.static:
            return 1.0
        return math.exp(-(self.a ** 2 + self.b ** 2) * self.viscosity() * time)

    # -- analytic fields on the whole lattice ---------------------------------
    def _grid(self):
        x, y = self.range()
        return np.meshgrid(x, y, indexing="ij")

    def _pressure_shape(self, X, Y):
        return self.A ** 2 * np.cos(2 * self.a * X) + self.B ** 2 * np.cos(2 * self.b * Y)

    def pressure(self, q: Quadrature, time: float = 0.0) -> np.ndarray:
        """taylor_green_vortex.jl:49-69"""
        X, Y = self._grid()
        P = -(1 / 4) * self.rho_0 * self.decay(time) ** 2 * self._pressure_shape(X, Y)
        return 1.0 + q.speed_of_sound_squared * self.u_max ** 2 * P

    def density(self, q: Quadrature, time: float = 0.0) -> np.ndarray:
        """taylor_green_vortex.jl:36-47: the pressure"""
        return self.pressure(q, time)

    def velocity(self, time: float = 0.0):
        """taylor_green_vortex.jl:71-85"""
        X, Y = self._grid()
        d = self.decay(time)
        return (d * (self.A * np.cos(self.a * X) * np.sin(self.b * Y)),
                d * (self.B * np.sin(self.a * X) * np.cos(self.b * Y)))

    def _gradient_shape(self, X, Y):
        a, b, A, B = self.a, self.b, self.A, self.B
        u_x = -a * A * np.sin(a * X) * np.sin(b * Y)
        v_y = -b * B * np.sin(a * X) * np.sin(b * Y)
        u_y = b * A * np.cos(a * X) * np.cos(b * Y)
        v_x = a * B * np.cos(a * X) * np.cos(b * Y)
        return u_x, v_x, u_y, v_y

    def deviatoric_tensor(self, q: Quadrature, time: float = 0.0):
        """-ν [2 a11, a12 + a21; a12 + a21, 2 a22] of the velocity gradient
        (problems.jl:35-50, taylor_green_vortex.jl:87-103): xx, xy, yx, yy."""
        X, Y = self._grid()
        d = self.decay(time)
        u_x, v_x, u_y, v_y = (d * g for g in self._gradient_shape(X, Y))
        nu = self.viscosity()
        off = -nu * (v_x + u_y)
        return -nu * (2 * u_x), off, off, -nu * (2 * v_y)

    # -- what TrackHydrodynamicErrors reads -----------------------------------
    def expected_fields(self, q: Quadrature, time: float) -> np.ndarray:
        """(NX, NY, 8): rho, u_x, u_y, p, sigma_xx, sigma_xy, sigma_yx, sigma_yy"""
        p = self.pressure(q, time)
        ux, uy = self.velocity(time)
        sxx, sxy, syx, syy = self.deviatoric_tensor(q, time)
        return np.stack([p, ux, uy, p, sxx, sxy, syx, syy], axis=-1)

    def expected_planes(self, q: Quadrature) -> np.ndarray:
        """The time-independent shapes: pressure and density as
        1 + factor * (A² cos 2ax + B² cos 2by), velocity and stress as
        factor * (their value at decay 1)."""
        X, Y = self._grid()
        shape = self._pressure_shape(X, Y)
        ux = self.A * np.cos(self.a * X) * np.sin(self.b * Y)
        uy = self.B * np.sin(self.a * X) * np.cos(self.b * Y)
        u_x, v_x, u_y, v_y = self._gradient_shape(X, Y)
        nu = self.viscosity()
        off = -nu * (v_x + u_y)
        return np.stack([shape, ux, uy, shape, -nu * (2 * u_x), off, off, -nu * (2 * v_y)],
                        axis=-1)

    def expected_affine(self, q: Quadrature, time: float):
        d = self.decay(time)
        pf = q.speed_of_sound_squared * self.u_max ** 2 * (-(1 / 4) * self.rho_0 * d ** 2)
        return (1.0, pf, 0.0, d, 0.0, d, 1.0, pf, 0.0, d, 0.0, d, 0.0, d, 0.0, d)

    # -- what the initial conditions read (problems.jl:98-106) -----------------
    def lattice_fields(self, q: Quadrature):
        """lattice_density, lattice_velocity = u_max velocity and
        lattice_temperature = pressure / density, as (NX, NY) planes."""
        p = self.pressure(q, 0.0)
        rho = self.density(q, 0.0)
        ux, uy = self.velocity(0.0)
        return rho, self.u_max * ux, self.u_max * uy, p / rho

    def velocity_gradient(self, time: float = 0.0) -> np.ndarray:
        """decay [u_x v_x; u_y v_y] (taylor_green_vortex.jl:87-104) as
        (NX, NY, 2, 2), G[..., a, b]."""
        X, Y = self._grid()
        d = self.decay(time)
        u_x, v_x, u_y, v_y = (d * g for g in self._gradient_shape(X, Y))
        return np.stack([np.stack([u_x, v_x], axis=-1), np.stack([u_y, v_y], axis=-1)], axis=-2)

    def offequilibrium_inputs(self):
        """The generic off-equilibrium start (analytical_offequilibrium.jl:10-48):
        gradient u_max^2 velocity_gradient and factor = the domain area."""
        from .initial_conditions import INIT_OFFEQ
        return (INIT_OFFEQ, self.u_max ** 2 * self.velocity_gradient(0.0),
                self.domain_size[0] * self.domain_size[1])

    # -- initial populations (AnalyticalEquilibrium, problems.jl:136-143) -------
    def initial_populations(self, q: Quadrature) -> np.ndarray:
        rho = self.density(q, 0.0)
        ux, uy = self.velocity(0.0)
        return equilibrium_populations(q, rho, self.u_max * ux, self.u_max * uy)

    def lattice_velocity_field(self) -> np.ndarray:
        """lattice_velocity = u_max velocity(x, y, 0) (problems.jl:91-92), (NX, NY, 2)"""
        ux, uy = self.velocity(0.0)
        return np.stack([self.u_max * ux, self.u_max * uy], axis=-1)

    def lattice_force(self):
        """u_max Δt force(x, y) = u_max Δt 2 ν_dimensionless u(x, y, 0) when
        static (taylor_green_vortex.jl:110-114, problems.jl:98-99), as
        (NX, NY, 2); None for the decaying vortex."""
        if not self.static:
            return None
        ux, uy = self.velocity(0.0)
        k = self.u_max * self.delta_t() * 2 * self.viscosity()
        return np.stack([k * ux, k * uy], axis=-1)


class TGV:
    """TGV(q, τ, scale = 2, NX = 16 scale, NY = NX, u_max = 0.02 / scale)
    (problems/second_order_convergence.jl:1-131): the Taylor-Green vortex of
    the convergence study, in lattice units (u_max = 1, domain (1, 1), node
    coordinates x NX), with its own equilibrium + off-equilibrium start
    (analytical_offequilibrium.jl:50-87)."""

    kind = "tgv"

    def __init__(self, q: Quadrature, τ: float, scale: int = 2, NX: int = None, NY: int = None,
                 u_max: float = None):
        self.q = q
        self.NX = 16 * scale if NX is None else int(NX)
        self.NY = self.NX if NY is None else int(NY)
        self.u_0 = 0.02 / scale if u_max is None else float(u_max)
        self.u_max = 1.0
        self.rho_0 = 1.0
        self.τ = float(τ)
        self.ν = (self.τ - 0.5) / q.speed_of_sound_squared
        self.domain_size = (1.0, 1.0)

    def lattice_viscosity(self) -> float:
        return self.ν

    def _grid(self):
        dx, dy = 1.0 / self.NX, 1.0 / self.NY
        x = np.linspace(dx / 2, 1.0 - dx / 2, self.NX) * self.NX
        y = np.linspace(dy / 2, 1.0 - dy / 2, self.NY) * self.NY
        return np.meshgrid(x, y, indexing="ij")

    def _k(self):
        return 2 * math.pi / self.NX, 2 * math.pi / self.NY

    def _td(self):
        kx, ky = self._k()
        return 1.0 / (self.ν * (kx * kx + ky * ky))

    def density(self, q: Quadrature, t: float = 0.0) -> np.ndarray:
        X, Y = self._grid()
        kx, ky = self._k()
        shape = (ky / kx) * np.cos(2 * kx * X) + (kx / ky) * np.cos(2 * ky * Y)
        return self.rho_0 * (1.0 - q.speed_of_sound_squared * (self.u_0 ** 2 / 4) * shape
                             * np.exp(-2 * t / self._td()))

    def pressure(self, q: Quadrature, t: float = 0.0) -> np.ndarray:
        X, Y = self._grid()
        kx, ky = self._k()
        shape = (ky / kx) * np.cos(2 * kx * X) + (kx / ky) * np.cos(2 * ky * Y)
        return self.rho_0 - (q.speed_of_sound_squared * 1.0 * (self.u_0 ** 2 / 4) * shape
                             * np.exp(-2 * t / self._td()))

    def velocity(self, t: float = 0.0):
        X, Y = self._grid()
        kx, ky = self._k()
        d = self.u_0 * np.exp(-t / self._td())
        return (d * (-math.sqrt(ky / kx) * np.cos(kx * X) * np.sin(ky * Y)),
                d * (math.sqrt(kx / ky) * np.sin(kx * X) * np.cos(ky * Y)))

    def velocity_gradient(self, t: float = 0.0) -> np.ndarray:
        X, Y = self._grid()
        kx, ky = self._k()
        u_x = math.sqrt(ky * kx) * np.sin(kx * X) * np.sin(ky * Y)
        v_y = -math.sqrt(ky * kx) * np.sin(kx * X) * np.sin(ky * Y)
        u_y = -math.sqrt(ky ** 3 / kx) * np.cos(kx * X) * np.cos(ky * Y)
        v_x = math.sqrt(kx ** 3 / ky) * np.cos(kx * X) * np.cos(ky * Y)
        d = np.exp(-t / self._td()) * self.u_0
        return d * np.stack([np.stack([u_x, v_x], axis=-1), np.stack([u_y, v_y], axis=-1)],
                            axis=-2)

    def lattice_fields(self, q: Quadrature):
        rho = self.density(q)
        ux, uy = self.velocity()
        return rho, ux, uy, self.pressure(q) / rho

    def offequilibrium_inputs(self):
        from .initial_conditions import INIT_OFFEQ_TGV
        return INIT_OFFEQ_TGV, self.velocity_gradient(0.0), 1.0

    def lattice_velocity_field(self) -> np.ndarray:
        ux, uy = self.velocity()
        return np.stack([ux, uy], axis=-1)
