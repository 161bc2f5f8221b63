"""Shared inputs for the parity tests: seeded populations, BC lists in both the
oracle's dict form and the lbmhip object form, and analytic flow fields."""
from __future__ import annotations

import numpy as np

QIDS = [4, 5, 9, 13, 17, 21, 37]


def random_populations(qid: int, nx: int, ny: int, seed: int = 0, u_amp: float = 0.03,
                       noise: float = 2e-3) -> np.ndarray:
    """Equilibrium of a random (rho, u) field plus a random non-equilibrium part."""
    from oracle import lbm_numpy as ln
    q = ln.by_id(qid)
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.01 * rng.standard_normal((nx, ny))
    ux = u_amp * rng.standard_normal((nx, ny))
    uy = u_amp * rng.standard_normal((nx, ny))
    f = ln.equilibrium_field(q, rho, ux, uy)
    f = f + noise * q.w * rng.standard_normal(f.shape)
    return np.asfortranarray(f)


# ---- boundary-condition lists -------------------------------------------------
def bc_dicts(name: str, nx: int, ny: int, u_wall=(0.05, 0.0)):
    full = dict(xs=(1, nx), ys=(1, ny))
    if name == "periodic":
        return []
    if name == "poiseuille":   # poiseuille.jl:93-96
        return [dict(kind=0, direction=0, **full), dict(kind=0, direction=2, **full)]
    if name == "couette":      # couette_flow.jl:82-85
        return [dict(kind=0, direction=2, **full),
                dict(kind=1, direction=0, u=u_wall, rho=1.0, **full)]
    if name == "cavity":       # lid_driven_cavity.jl:61-66
        return [dict(kind=0, direction=1, **full), dict(kind=0, direction=2, **full),
                dict(kind=0, direction=3, **full),
                dict(kind=1, direction=0, u=u_wall, rho=1.0, **full)]
    if name == "partial":      # ranges narrower than the lattice, list order matters
        lo_y, lo_x = min(2, ny), min(2, nx)
        return [dict(kind=0, direction=3, xs=(1, nx), ys=(lo_y, max(lo_y, ny - 1))),
                dict(kind=0, direction=0, xs=(lo_x, max(lo_x, nx - 1)), ys=(1, ny)),
                dict(kind=0, direction=1, xs=(1, nx), ys=(1, ny)),
                dict(kind=0, direction=2, xs=(1, nx), ys=(1, ny))]
    raise ValueError(name)


BC_NAMES = ["periodic", "poiseuille", "couette", "cavity", "partial"]


def to_lbmhip(bcs):
    import lbmhip as L
    dirs = {0: L.North(), 1: L.East(), 2: L.South(), 3: L.West()}
    out = []
    for b in bcs:
        if b["kind"] == 0:
            out.append(L.BounceBack(dirs[b["direction"]], b["xs"], b["ys"]))
        else:
            out.append(L.MovingWall(dirs[b["direction"]], b["xs"], b["ys"], b["u"],
                                    b.get("rho", 1.0)))
    return out


def force_for(kind: str, nx: int, ny: int, seed: int = 5):
    if kind == "none":
        return None
    if kind == "const":
        return np.array([1e-5, -3e-6])
    rng = np.random.default_rng(seed)
    return np.asfortranarray(1e-5 * rng.standard_normal((nx, ny, 2)))


# ---- analytic initial fields ----------------------------------------------------
def shear_wave_fields(n: int, amp: float = 0.01):
    """C1: rho = 1, u_x = amp sin(2 pi y / n), u_y = 0 (y = node index)."""
    y = np.arange(n)
    ux = np.tile(amp * np.sin(2 * np.pi * y / n), (n, 1))
    return np.ones((n, n)), ux, np.zeros((n, n))


def taylor_green_fields(n: int, u0: float = 0.02):
    """C2: u = u0 (sin x cos y, -cos x sin y) on cell centres of [0, 2 pi)^2,
    rho = 1 - (3 u0^2 / 4)(cos 2x + cos 2y)  (p = rho / 3)."""
    c = (np.arange(n) + 0.5) * (2 * np.pi / n)
    X, Y = np.meshgrid(c, c, indexing="ij")
    ux = u0 * np.sin(X) * np.cos(Y)
    uy = -u0 * np.cos(X) * np.sin(Y)
    rho = 1.0 - (3.0 * u0 * u0 / 4.0) * (np.cos(2 * X) + np.cos(2 * Y))
    return rho, ux, uy


def taylor_green_vortex_pops(q, n: int, u0: float = 0.02):
    import lbmhip as L
    rho, ux, uy = taylor_green_fields(n, u0)
    return L.equilibrium_populations(q, rho, ux, uy)
