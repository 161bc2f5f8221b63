"""Vectorised NumPy restatement of LatticeBoltzmann.jl's hot path.

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

This module deliberately shares no code or structure with ``lbm_oracle.c``:
quadrature tables are re-transcribed from the reference by velocity group,
streaming is ``np.roll`` per plane, boundary conditions are boolean masks over
the whole lattice, and the equilibrium is also available as a tensor-Hermite
sum.  Agreement between the two restatements to rounding is what pins the
C oracle's indexing (layout, opposite map, BC predicates, wrap rule); the
physics is pinned separately by the closed-form and analytic checks in tests.

Arrays are float64 ``(nx, ny, Q)``; the reference's (src/initial_conditions.jl:13).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class Quadrature:
    name: str
    c: np.ndarray          # (2, Q) int
    w: np.ndarray          # (Q,)
    s: float               # speed_of_sound_squared (= 1/c_s^2)
    order: int
    feq_order: int

    @property
    def Q(self) -> int:
        return self.w.size

    def opposite(self) -> np.ndarray:
        """Index j with c_j = -c_i, found by search (not by the reference's rule)."""
        Q = self.Q
        out = np.empty(Q, dtype=np.int64)
        for i in range(Q):
            hits = [j for j in range(Q) if self.c[0, j] == -self.c[0, i]
                    and self.c[1, j] == -self.c[1, i]]
            out[i] = hits[0]
        return out


def _quad(name):
    # Velocity groups transcribed from src/quadratures/D2Q*.jl.
    if name == "D2Q4":
        c = [(1, 0), (0, 1), (-1, 0), (0, -1)]
        w = [0.25] * 4
        return Quadrature(name, np.array(c).T, np.array(w), 2.0, 3, 1)
    if name == "D2Q5":
        c = [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1)]
        w = [4 / 6] + [1 / 12] * 4
        return Quadrature(name, np.array(c).T, np.array(w), 6.0, 3, 1)
    if name == "D2Q9":
        c = [(0, 0), (-1, 1), (-1, 0), (-1, -1), (0, -1), (1, -1), (1, 0), (1, 1), (0, 1)]
        w = [4 / 9] + [1 / 36, 1 / 9] * 4
        return Quadrature(name, np.array(c).T, np.array(w), 3.0, 5, 2)
    if name == "D2Q13":
        c = [(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1), (1, -1), (-1, 1),
             (2, 0), (-2, 0), (0, -2), (0, 2)]
        w = [3 / 8] + [1 / 12] * 4 + [1 / 16] * 4 + [1 / 96] * 4
        return Quadrature(name, np.array(c).T, np.array(w), 2.0, 5, 2)
    if name == "D2Q17":
        sq = math.sqrt(193)
        ws = [(575 + 193 * sq) / 8100, (3355 - 91 * sq) / 18000, (655 + 17 * sq) / 27000,
              (685 - 49 * sq) / 54000, (1445 - 101 * sq) / 162000]
        c = [(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1), (1, -1), (-1, 1),
             (2, 2), (-2, -2), (2, -2), (-2, 2), (3, 0), (-3, 0), (0, 3), (0, -3)]
        w = [ws[0]] + [ws[1]] * 4 + [ws[2]] * 4 + [ws[3]] * 4 + [ws[4]] * 4
        return Quadrature(name, np.array(c).T, np.array(w), (125 + 5 * math.sqrt(193)) / 72,
                          7, 3)
    if name == "D2Q21":
        ws = [91 / 324, 1 / 12, 2 / 27, 7 / 360, 1 / 432, 1 / 1620, 0.0]
        c = [(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1), (1, -1), (-1, 1),
             (2, 0), (-2, 0), (0, 2), (0, -2), (2, 2), (-2, -2), (2, -2), (-2, 2),
             (3, 0), (-3, 0), (0, 3), (0, -3), (3, 3), (-3, -3), (-3, 3), (3, -3)]
        w = [ws[0]] + sum(([x] * 4 for x in ws[1:]), [])
        return Quadrature(name, np.array(c).T, np.array(w), 1.5, 7, 3)
    if name == "D2Q37":
        g = [0.23315066913235250228650, 0.10730609154221900241246, 0.05766785988879488203006,
             0.01420821615845075026469, 0.00535304900051377523273, 0.00101193759267357547541,
             0.00024530102775771734547, 0.00028341425299419821740]
        c = [(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1), (1, -1), (-1, 1),
             (2, 0), (-2, 0), (0, 2), (0, -2),
             (2, 1), (-2, -1), (-2, 1), (2, -1), (1, 2), (-1, -2), (1, -2), (-1, 2),
             (2, 2), (-2, -2), (2, -2), (-2, 2), (3, 0), (-3, 0), (0, 3), (0, -3),
             (3, 1), (-3, -1), (3, -1), (-3, 1), (1, 3), (-1, -3), (-1, 3), (1, -3)]
        w = [g[0]] + [g[1]] * 4 + [g[2]] * 4 + [g[3]] * 4 + [g[4]] * 8 + [g[5]] * 4 \
            + [g[6]] * 4 + [g[7]] * 8
        r = 1.19697977039307435897239
        return Quadrature(name, np.array(c).T, np.array(w), r * r, 9, 4)
    raise ValueError(name)


QUADRATURES = {n: _quad(n) for n in
               ["D2Q4", "D2Q5", "D2Q9", "D2Q13", "D2Q17", "D2Q21", "D2Q37"]}


def by_id(qid: int) -> Quadrature:
    return QUADRATURES[f"D2Q{qid}"]


# ---------------------------------------------------------------------------
# Hermite tensors (src/hermite_polynomials.jl:47-82), built from the
# Rodrigues-style recurrence H_{n+1} = xi (x) H_n - cs * sym(delta (x) H_{n-1})
# rather than the reference's closed forms, as an independent formulation.
# ---------------------------------------------------------------------------
def hermite_tensor(n: int, xi, cs: float) -> np.ndarray:
    xi = np.asarray(xi, dtype=np.float64)
    I = np.eye(2)
    if n == 0:
        return np.array(1.0)
    if n == 1:
        return xi.copy()
    if n == 2:
        return np.einsum("a,b->ab", xi, xi) - cs * I
    if n == 3:
        t = np.einsum("a,b,c->abc", xi, xi, xi)
        d = (np.einsum("a,bc->abc", xi, I) + np.einsum("b,ac->abc", xi, I)
             + np.einsum("c,ab->abc", xi, I))
        return t - cs * d
    if n == 4:
        t = np.einsum("a,b,c,d->abcd", xi, xi, xi, xi)
        xx = np.einsum("a,b->ab", xi, xi)
        d2 = (np.einsum("ab,cd->abcd", xx, I) + np.einsum("ac,bd->abcd", xx, I)
              + np.einsum("ad,bc->abcd", xx, I) + np.einsum("bc,ad->abcd", xx, I)
              + np.einsum("bd,ac->abcd", xx, I) + np.einsum("cd,ab->abcd", xx, I))
        d0 = (np.einsum("ab,cd->abcd", I, I) + np.einsum("ac,bd->abcd", I, I)
              + np.einsum("ad,bc->abcd", I, I))
        return t - cs * d2 + cs * cs * d0
    raise ValueError(n)


def hermite_equilibrium(q: Quadrature, rho: float, u) -> np.ndarray:
    """f_i = w_i (rho + sum_n a_eq^(n) : H_n(c_i) / (n! cs^n)), T = 1
    (velocity_distribution_function/hermite.jl:10-33, 37-77)."""
    u = np.asarray(u, dtype=np.float64)
    cs = 1.0 / q.s
    N = q.order // 2
    out = np.empty(q.Q)
    aeq = {1: rho * u,
           2: rho * np.einsum("a,b->ab", u, u),
           3: rho * np.einsum("a,b,c->abc", u, u, u),
           4: rho * np.einsum("a,b,c,d->abcd", u, u, u, u)}
    for i in range(q.Q):
        tot = 0.0
        for n in range(1, N + 1):
            H = hermite_tensor(n, q.c[:, i], cs)
            tot += float(np.sum(aeq[n] * H)) / (math.factorial(n) * cs ** n)
        out[i] = q.w[i] * (rho + tot)
    return out


# ---------------------------------------------------------------------------
# Closed-form equilibrium over whole fields (maxwell_boltzmann_equilibrium.jl,
# velocity_distribution_function/quadratures.jl); written as a sum of
# Hermite orders in powers of (s*ud) and (s*usq).
# ---------------------------------------------------------------------------
def equilibrium_field(q: Quadrature, rho, ux, uy) -> np.ndarray:
    s = q.s
    rho = np.asarray(rho, dtype=np.float64)
    usq = ux * ux + uy * uy
    out = np.empty(rho.shape + (q.Q,), dtype=np.result_type(rho, ux, uy, 1.0))
    for i in range(q.Q):
        ud = q.c[0, i] * ux + q.c[1, i] * uy
        x = s * ud
        y = s * usq
        poly = 1.0 + x
        if q.feq_order >= 2:
            poly = poly + 0.5 * (x * x - y)
        if q.feq_order >= 3:
            poly = poly + (x * (x * x - 3.0 * y)) / 6.0
        if q.feq_order >= 4:
            poly = poly + (x ** 4 - 6.0 * y * x * x + 3.0 * y * y) / 24.0
        out[..., i] = rho * q.w[i] * poly
    return out


def moments_field(q: Quadrature, f):
    rho = f.sum(axis=2)
    ux = (f * q.c[0]).sum(axis=2) / rho
    uy = (f * q.c[1]).sum(axis=2) / rho
    return rho, ux, uy


def _force_field(force, nx, ny):
    if force is None:
        return None
    F = np.asarray(force, dtype=np.float64)
    if F.size == 2:
        return np.broadcast_to(F.reshape(1, 1, 2), (nx, ny, 2))
    return F.reshape(nx, ny, 2, order="F")


def collide_srt(q, f, tau, force=None):
    nx, ny, _ = f.shape
    rho, ux, uy = moments_field(q, f)
    F = _force_field(force, nx, ny)
    if F is not None:
        ux = ux + tau * F[..., 0]
        uy = uy + tau * F[..., 1]
    feq = equilibrium_field(q, rho, ux, uy)
    return f - (f - feq) / tau


def collide_trt(q, f, tau_s, tau_a, force=None):
    nx, ny, _ = f.shape
    rho, ux, uy = moments_field(q, f)
    F = _force_field(force, nx, ny)
    if F is not None:
        ux = ux + tau_a * F[..., 0]
        uy = uy + tau_a * F[..., 1]
    feq = equilibrium_field(q, rho, ux, uy)
    opp = q.opposite()
    neq = f - feq
    sym = 0.5 * (neq + neq[..., opp])
    asym = 0.5 * (neq - neq[..., opp])
    return f - sym / tau_s - asym / tau_a


def collide_mrt(q, f, taus, force=None):
    """Regularised projection of src/collision_models/mrt.jl:56-118 written as
    tensor contractions: a_n = sum_i f_i H_n(c_i), relax, reconstruct."""
    nx, ny, Q = f.shape
    cs = 1.0 / q.s
    N = q.order // 2
    rho, ux, uy = moments_field(q, f)
    F = _force_field(force, nx, ny)
    if F is not None:
        ux = ux + taus[1] * F[..., 0]
        uy = uy + taus[1] * F[..., 1]
    u = np.stack([ux, uy], axis=-1)
    out = rho[..., None] * (1.0 + q.s * np.einsum("xyd,di->xyi", u, q.c.astype(float)))
    out = out * q.w
    for n in range(2, N + 1):
        Hn = np.stack([hermite_tensor(n, q.c[:, i], cs).reshape(-1) for i in range(Q)])
        af = np.einsum("xyi,ik->xyk", f, Hn)
        if n == 2:
            ueq = np.einsum("xya,xyb->xyab", u, u)
        elif n == 3:
            ueq = np.einsum("xya,xyb,xyc->xyabc", u, u, u)
        else:
            ueq = np.einsum("xya,xyb,xyc,xyd->xyabcd", u, u, u, u)
        aeq = rho[..., None] * ueq.reshape(nx, ny, -1)
        t = taus[n - 1]
        ac = af - (af - aeq) / t
        out = out + q.w * (q.s ** n) * np.einsum("xyk,ik->xyi", ac, Hn) / math.factorial(n)
    return out


def stream(q, f):
    """Pull streaming = per-plane periodic roll by c_i (src/stream.jl:19-30)."""
    out = np.empty_like(f)
    for i in range(q.Q):
        out[:, :, i] = np.roll(f[:, :, i], shift=(int(q.c[0, i]), int(q.c[1, i])), axis=(0, 1))
    return out


def apply_bcs(q, bcs, fnew, fold):
    """Ordered BC list with boolean masks (boundary_conditions.jl:6-20,
    bounce_back.jl, moving_wall.jl).  bcs: dicts as in oracle.clib.make_bcs."""
    nx, ny, Q = fnew.shape
    out = fnew.copy()
    opp = q.opposite()
    X, Y = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), indexing="ij")
    for bc in bcs:
        if bc["kind"] == 0:
            (x0, x1), (y0, y1) = bc["xs"], bc["ys"]
            inrange = (X >= x0) & (X <= x1) & (Y >= y0) & (Y <= y1)
        else:
            inrange = np.ones((nx, ny), dtype=bool)
        for i in range(Q):
            srcx = X - q.c[0, i]
            srcy = Y - q.c[1, i]
            d = bc["direction"]
            beyond = {0: srcy > ny, 2: srcy < 1, 1: srcx > nx, 3: srcx < 1}[d]
            m = inrange & beyond
            val = fold[:, :, opp[i]]
            if bc["kind"] == 1:
                ux, uy = bc.get("u", (0.0, 0.0))
                rw = bc.get("rho", 1.0)
                val = val + 2.0 * q.w[i] * q.s * rw * (ux * q.c[0, i] + uy * q.c[1, i])
            out[:, :, i] = np.where(m, val, out[:, :, i])
    return out


def simulate(q, f_stream, collision, taus, nsteps, force=None, bcs=()):
    """Returns (f_stream, f_collision) after ``nsteps`` reference time steps."""
    S = f_stream.copy()
    C = None
    for _ in range(nsteps):
        if collision == 0:
            C = collide_srt(q, S, taus[0], force)
        elif collision == 1:
            C = collide_trt(q, S, taus[0], taus[1], force)
        else:
            C = collide_mrt(q, S, taus, force)
        S = apply_bcs(q, list(bcs), stream(q, C), C)
    return S, C


def deviatoric_field(q, f, tau):
    """sigma = dev((sum_i f_i H2_i - rho u u) / (1 + 1/(2 tau))) (moments.jl:81-96)."""
    cs = 1.0 / q.s
    rho, ux, uy = moments_field(q, f)
    cc = q.c.astype(float)
    a = np.einsum("xyi,ai,bi->xyab", f, cc, cc) - cs * rho[..., None, None] * np.eye(2)
    u = np.stack([ux, uy], axis=-1)
    a = a - rho[..., None, None] * np.einsum("xya,xyb->xyab", u, u)
    sig = a / (1.0 + 1.0 / (2.0 * tau))
    tr = 0.5 * (sig[..., 0, 0] + sig[..., 1, 1])
    return sig[..., 0, 0] - tr, sig[..., 0, 1], sig[..., 1, 1] - tr


def track_errors(q, f, expected, affine, tau, u_max, delta=1.0):
    """One TrackHydrodynamicErrors row (track_hydrodynamic_errors.jl:52-222) as
    whole-lattice tensor algebra: the second Hermite moment, the pressure as
    the trace of P = a_2 - rho (u u - I), sigma from the deviatoric part.
    expected: (nx, ny, 8) planes rho, ux, uy, p, sxx, sxy, syx, syy."""
    a = np.asarray(affine, dtype=np.float64).reshape(8, 2)
    e = a[:, 0] + a[:, 1] * np.asarray(expected, dtype=np.float64)
    cs = 1.0 / q.s
    rho, ux, uy = moments_field(q, f)
    cc = q.c.astype(float)
    abar = np.einsum("xyi,ai,bi->xyab", f, cc, cc) - cs * rho[..., None, None] * np.eye(2)
    u = np.stack([ux, uy], axis=-1)
    uu = np.einsum("xya,xyb->xyab", u, u)
    aeq = rho[..., None, None] * uu
    k = 1.0 / (2.0 * tau)
    a2 = (abar + k * aeq) / (1.0 + k)
    P = a2 - rho[..., None, None] * (uu - np.eye(2))
    p = np.trace(P, axis1=2, axis2=3) / 2.0
    sig = (abar - aeq) / (1.0 + k)
    sig = sig - np.eye(2) * (np.trace(sig, axis1=2, axis2=3) / 2.0)[..., None, None]
    sig = sig / u_max ** 2
    d = u / u_max
    ex_s = {(0, 0): e[..., 4], (0, 1): e[..., 5], (1, 0): e[..., 6], (1, 1): e[..., 7]}

    def rel(err, ref):
        return math.sqrt(np.sum(err ** 2) / np.sum(ref ** 2))

    out = [math.sqrt(np.sum((rho - e[..., 0]) ** 2)),
           math.sqrt(np.sum((d[..., 0] - e[..., 1]) ** 2 + (d[..., 1] - e[..., 2]) ** 2)
                     / np.sum(e[..., 1] ** 2 + e[..., 2] ** 2)),
           rel(p - e[..., 3], e[..., 3])]
    for ab in ((0, 0), (0, 1), (1, 1), (1, 0)):
        out.append(rel(ex_s[ab] - sig[..., ab[0], ab[1]], ex_s[ab]))
    out += [delta * np.sum(rho), delta * np.sum(rho * (d[..., 0] + d[..., 1])),
            delta * np.sum(rho * (d[..., 0] ** 2 + d[..., 1] ** 2)),
            float(np.sum(~np.all(np.isfinite(f), axis=2)))]
    return np.array(out)


def velocity_stats(q, f, u_old=None):
    """Stopping-criteria sums (stopping_criteria.jl:17-99): mean ux, mean uy,
    sum |u - u_old|^2, sum |u_old|^2, non-finite nodes."""
    rho, ux, uy = moments_field(q, f)
    out = [ux.mean(), uy.mean(), 0.0, 0.0, float(np.sum(~np.all(np.isfinite(f), axis=2)))]
    if u_old is not None:
        out[2] = float(np.sum((ux - u_old[..., 0]) ** 2 + (uy - u_old[..., 1]) ** 2))
        out[3] = float(np.sum(u_old ** 2))
    return np.array(out)


def mei_nonlinear(q, u):
    """The nonlinear equilibrium term of the iterative initialization
    (iterative_initialization.jl:12-40) as tensor algebra:
    w_i (s c.u + (s^2 (c.u)^2 - s |u|^2) / 2)."""
    cu = np.einsum("xyd,di->xyi", np.asarray(u, dtype=np.float64), q.c.astype(float))
    usq = np.sum(np.asarray(u) ** 2, axis=2)[..., None]
    return q.w * (q.s * cu + 0.5 * (q.s ** 2 * cu ** 2 - q.s * usq))


def collide_mei(q, f, tau, N):
    rho = f.sum(axis=2)[..., None]
    return (1.0 - 1.0 / tau) * f + (1.0 / tau) * (q.w * rho + N)
