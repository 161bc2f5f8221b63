"""Host-side initial populations (src/initial_conditions.jl:7-22, 182-191).

``equilibrium_populations(q, ρ, ux, uy)`` fills f[x, y, :] with the truncated
Hermite equilibrium of each node, the AnalyticalEquilibrium strategy
(initial_conditions/analytical_equilibrium.jl).  The reference evaluates it
with the generic Hermite sum (velocity_distribution_function/hermite.jl:10-33);
the closed form used here is the same polynomial and agrees to rounding.
Initialization is one-shot host work, outside the timed step.
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
