"""Reverse-mode checker for the differentiable collide-stream run.

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

The reference has no adjoint: it is differentiated only by running it.  This
module computes the gradients of

    L(u0, tau) = 1/2 sum_x |u(S_T)[x] - u_target[x]|^2,
    S_0 = feq(rho0, u0),  S_{t+1} = stream(collide(S_t))     (periodic, unforced)

with explicit per-node Jacobian matrices, written without reference to the
HIP kernels' formulation:

* d feq / d u by complex-step differentiation of ``lbm_numpy.equilibrium_field``
  (exact to rounding, no hand-derived polynomial derivative);
* the collision Jacobian as a dense Q x Q matrix per node:
  SRT  J = (1 - om) I + om G,
  TRT  J = I - (om_s P+ + om_a P-) (I - G),  P+- = (I +- R)/2, R the opposite map,
  G_ij = d feq_i / d f_j = feq_i / rho + sum_d d feq_i/d u_d (c_jd - u_d) / rho;
* the transposed stream as the inverse ``np.roll``.

Its parity is pinned by central finite differences of the C restatement's
FP64 forward run (tests/test_oracle_adjoint.py); the GPU reverse pass is then
checked against it node by node.
"""
from __future__ import annotations

import numpy as np

from . import lbm_numpy as ln

_H = 1e-30  # complex-step size


def _dfeq_du(q, rho, ux, uy):
    """(nx, ny, Q, 2): d feq_i / d u_d at (rho, u), by complex step."""
    gx = np.imag(ln.equilibrium_field(q, rho, ux + 1j * _H, uy + 0j)) / _H
    gy = np.imag(ln.equilibrium_field(q, rho, ux + 0j, uy + 1j * _H)) / _H
    return np.stack([gx, gy], axis=-1)


def _G(q, f):
    rho, ux, uy = ln.moments_field(q, f)
    feq = ln.equilibrium_field(q, rho, ux, uy)
    D = _dfeq_du(q, rho, ux, uy)                                   # (nx, ny, Q, 2)
    u = np.stack([ux, uy], axis=-1)                                # (nx, ny, 2)
    dudf = (q.c.T[None, None, :, :] - u[:, :, None, :]) / rho[:, :, None, None]  # (nx,ny,Q_j,2)
    G = feq[..., :, None] / rho[:, :, None, None] + np.einsum("xyid,xyjd->xyij", D, dudf)
    return G, feq


def _relax_matrix(q, coll, taus):
    """om * I (SRT) or om_s P+ + om_a P- (TRT), Q x Q."""
    Q = q.Q
    if coll == 0:
        return np.eye(Q) / taus[0]
    R = np.zeros((Q, Q))
    R[np.arange(Q), q.opposite()] = 1.0
    Pp, Pm = 0.5 * (np.eye(Q) + R), 0.5 * (np.eye(Q) - R)
    return Pp / taus[0] + Pm / taus[1]


def _collide(q, f, coll, taus):
    return ln.collide_srt(q, f, taus[0]) if coll == 0 else ln.collide_trt(q, f, taus[0], taus[1])


def forward(q, u0, taus, coll, nsteps, rho0=1.0):
    """States S_0 .. S_T as a list of (nx, ny, Q) arrays."""
    nx, ny = u0.shape[:2]
    S = [ln.equilibrium_field(q, np.full((nx, ny), rho0), u0[..., 0], u0[..., 1])]
    for _ in range(nsteps):
        S.append(ln.stream(q, _collide(q, S[-1], coll, taus)))
    return S


def loss(q, u0, u_target, taus, coll, nsteps, rho0=1.0):
    S = forward(q, u0, taus, coll, nsteps, rho0)
    _, ux, uy = ln.moments_field(q, S[-1])
    return 0.5 * float(np.sum((ux - u_target[..., 0]) ** 2 + (uy - u_target[..., 1]) ** 2))


def gradient(q, u0, u_target, taus, coll, nsteps, rho0=1.0):
    """(L, dL/du0 (nx, ny, 2), dL/dtau as a list: [dtau] (SRT) or [dtau_s, dtau_a])."""
    S = forward(q, u0, taus, coll, nsteps, rho0)
    rho, ux, uy = ln.moments_field(q, S[-1])
    d = np.stack([ux - u_target[..., 0], uy - u_target[..., 1]], axis=-1)
    L = 0.5 * float(np.sum(d * d))
    u = np.stack([ux, uy], axis=-1)
    lam = np.einsum("xyd,xyjd->xyj", d, (q.c.T[None, None] - u[:, :, None, :])) / rho[..., None]
    W = _relax_matrix(q, coll, taus)
    I = np.eye(q.Q)
    if coll == 0:
        projs = [I]
    else:
        R = np.zeros((q.Q, q.Q))
        R[np.arange(q.Q), q.opposite()] = 1.0
        projs = [0.5 * (I + R), 0.5 * (I - R)]
    gt = np.zeros(len(projs))
    for t in range(nsteps - 1, -1, -1):
        mu = np.empty_like(lam)
        for i in range(q.Q):   # transposed stream: undo the roll by c_i
            mu[..., i] = np.roll(lam[..., i], shift=(-int(q.c[0, i]), -int(q.c[1, i])), axis=(0, 1))
        G, feq = _G(q, S[t])
        neq = S[t] - feq
        for k, P in enumerate(projs):
            gt[k] += float(np.einsum("xyi,ij,xyj->", mu, P, neq)) / taus[k] ** 2
        # J = I - W (I - G);  J^T mu = mu - (I - G)^T W^T mu
        Wmu = np.einsum("ij,xyi->xyj", W, mu)               # (W^T mu)_j
        lam = mu - (Wmu - np.einsum("xyij,xyi->xyj", G, Wmu))
    D0 = _dfeq_du(q, np.full(u0.shape[:2], rho0), u0[..., 0], u0[..., 1])
    grad_u0 = np.einsum("xyi,xyid->xyd", lam, D0)
    return L, grad_u0, list(gt)
