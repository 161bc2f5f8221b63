"""Reverse-mode gradients through the collide-stream time step.

The differentiated program is the reference's plain loop on a periodic,
unforced lattice: equilibrium initial condition ``S_0 = feq(rho0, u0)``
(src/initial_conditions.jl:7-22), ``nsteps`` iterations of ``collide!`` then
``stream!`` (src/lattice_boltzmann_model.jl:60-70), and the loss

    L(u0, tau) = 1/2 * sum_x |u(S_T)[x] - u_target[x]|^2

with ``u`` the reference's ``velocity!`` moment (src/velocity_distribution_function/moments.jl:5-19).  The
gradients come from liblbm_hip's hand-written reverse pass (transposed
stream = push, transposed collision Jacobian per node) with checkpoints every
``checkpoint_every`` steps (default ceil(sqrt(nsteps))).  The reference has no
adjoint; :func:`rrule` has the shape of ChainRulesCore's ``rrule`` so that a
Julia binding can forward to it unchanged.

Velocity fields are float64 ``(nx, ny, 2)`` arrays, the layout of the
reference's per-node velocity vectors stacked over the lattice.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Optional, Tuple, Union

import numpy as np

from . import _lib
from .collision_models import SRT, TRT, CollisionModel
from .operators import NativeModel
from .quadratures import Quadrature


def _vfield(a, shape) -> np.ndarray:
    v = np.asfortranarray(a, dtype=np.float64)
    if v.shape != tuple(shape) + (2,):
        raise ValueError(f"velocity field must have shape {tuple(shape) + (2,)}, got {v.shape}")
    return v


class AdjointProblem:
    """A differentiable run: lattice shape, step count, storage and the
    collision model (SRT or TRT; its relaxation times may be changed between
    calls with :meth:`set_collision`)."""

    def __init__(self, q: Quadrature, cm: CollisionModel, shape: Tuple[int, int], nsteps: int,
                 storage: str = "f64", rho0: float = 1.0, checkpoint_every: int = 0):
        if not isinstance(cm, (SRT, TRT)):
            raise TypeError("the adjoint supports SRT and TRT collision models")
        if storage not in ("f64", "f32"):
            raise ValueError("adjoint storage must be 'f64' or 'f32'")
        self.q = q
        self.shape = (int(shape[0]), int(shape[1]))
        self.nsteps = int(nsteps)
        self.storage = storage
        self.rho0 = float(rho0)
        self._model = NativeModel(q)
        self._model.set_collision(cm)
        self.cm = cm
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().lbm_adjoint_create(
            ctypes.byref(h), self._model.handle, _lib.STORAGES[storage], self.shape[0],
            self.shape[1], self.rho0, self.nsteps, int(checkpoint_every)), "lbm_adjoint_create")
        self._h = h

    def set_collision(self, cm: CollisionModel) -> None:
        if not isinstance(cm, (SRT, TRT)):
            raise TypeError("the adjoint supports SRT and TRT collision models")
        self._model.set_collision(cm)
        self.cm = cm

    def info(self) -> dict:
        k, n = ctypes.c_int(), ctypes.c_int()
        b = ctypes.c_int64()
        _lib.check(_lib.lib().lbm_adjoint_info(self._h, ctypes.byref(k), ctypes.byref(n),
                                               ctypes.byref(b)), "lbm_adjoint_info")
        return {"checkpoint_every": k.value, "checkpoints": n.value, "bytes": b.value}

    def loss(self, u0, u_target, return_velocity: bool = False):
        """Forward pass only: L, and u(S_T) if ``return_velocity``."""
        a, t = _vfield(u0, self.shape), _vfield(u_target, self.shape)
        out = ctypes.c_double()
        uf = np.empty(self.shape + (2,), order="F") if return_velocity else None
        _lib.check(_lib.lib().lbm_adjoint_loss(self._h, _lib.dptr(a), _lib.dptr(t),
                                               ctypes.byref(out), _lib.dptr(uf)),
                   "lbm_adjoint_loss")
        return (out.value, uf) if return_velocity else out.value

    def gradient(self, u0, u_target) -> dict:
        """L, dL/du0 ((nx, ny, 2)) and dL/dtau (SRT: float; TRT: (dL/dtau_s,
        dL/dtau_a)), plus the device time of the whole pass in ms."""
        a, t = _vfield(u0, self.shape), _vfield(u_target, self.shape)
        out = ctypes.c_double()
        g = np.empty(self.shape + (2,), order="F")
        gt = np.zeros(2)
        ms = ctypes.c_double()
        _lib.check(_lib.lib().lbm_adjoint_gradient(self._h, _lib.dptr(a), _lib.dptr(t),
                                                   ctypes.byref(out), _lib.dptr(g),
                                                   _lib.dptr(gt), ctypes.byref(ms)),
                   "lbm_adjoint_gradient")
        grad_tau: Union[float, Tuple[float, float]] = (
            float(gt[0]) if isinstance(self.cm, SRT) else (float(gt[0]), float(gt[1])))
        return {"loss": out.value, "grad_u0": g, "grad_tau": grad_tau, "ms": ms.value}

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.lib().lbm_adjoint_destroy(h)
        self._h = None
        m = getattr(self, "_model", None)
        if m is not None:
            m.close()
            self._model = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def simulate_loss(q: Quadrature, u0, tau, u_target, nsteps: int, storage: str = "f64",
                  rho0: float = 1.0) -> float:
    """L(u0, tau) for SRT(tau) (``tau`` a number) or TRT with ``tau = (tau_s, tau_a)``."""
    cm = SRT(float(tau)) if np.ndim(tau) == 0 else TRT.fields(float(tau[0]), float(tau[1]))
    prob = AdjointProblem(q, cm, np.shape(u0)[:2], nsteps, storage=storage, rho0=rho0)
    try:
        return prob.loss(u0, u_target)
    finally:
        prob.close()


def rrule(f: Callable, q: Quadrature, u0, tau, u_target, nsteps: int, storage: str = "f64",
          rho0: float = 1.0, checkpoint_every: int = 0):
    """ChainRulesCore-style rule for :func:`simulate_loss`: returns
    ``(L, pullback)`` with ``pullback(dL) -> (NoTangent, NoTangent, dL * dL/du0,
    dL * dL/dtau, NoTangent, ...)`` (``None`` stands for NoTangent)."""
    if f is not simulate_loss:
        raise TypeError("rrule is defined for simulate_loss only")
    cm = SRT(float(tau)) if np.ndim(tau) == 0 else TRT.fields(float(tau[0]), float(tau[1]))
    prob = AdjointProblem(q, cm, np.shape(u0)[:2], nsteps, storage=storage, rho0=rho0,
                          checkpoint_every=checkpoint_every)
    try:
        res = prob.gradient(u0, u_target)
    finally:
        prob.close()
    gu, gt = res["grad_u0"], res["grad_tau"]

    def pullback(dL: float):
        s = float(dL)
        t = s * gt if np.ndim(gt) == 0 else (s * gt[0], s * gt[1])
        return (None, None, s * gu, t, None, None)

    return res["loss"], pullback


def divergence_free_field(n: Tuple[int, int], kmax: int, amplitude: float,
                          seed: Optional[int] = None) -> np.ndarray:
    """A seeded random periodic velocity field (nx, ny, 2) from a stream
    function of Fourier modes with 1 <= |k| <= kmax, scaled to max |u| =
    amplitude.  Its discrete divergence in Fourier space is zero."""
    nx, ny = int(n[0]), int(n[1])
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(nx, 1.0 / nx)[:, None]
    ky = np.fft.fftfreq(ny, 1.0 / ny)[None, :]
    kk = np.sqrt(kx * kx + ky * ky)
    mask = (kk >= 1) & (kk <= kmax)
    psi_hat = (rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))) * mask
    psi_hat = np.fft.fft2(np.real(np.fft.ifft2(psi_hat)))   # real stream function
    # u = (d psi / dy, -d psi / dx), spectral derivatives
    ux = np.real(np.fft.ifft2(1j * (2 * np.pi * ky / ny) * psi_hat))
    uy = np.real(np.fft.ifft2(-1j * (2 * np.pi * kx / nx) * psi_hat))
    u = np.stack([ux, uy], axis=-1)
    peak = np.sqrt((u * u).sum(axis=-1)).max()
    return np.asfortranarray(u * (amplitude / peak))
