"""lbmhip -- the collide-stream step of LatticeBoltzmann.jl on MI355X (gfx950).

Host-side mirror of the reference's Julia API over the C ABI of
liblbm_hip.so (include/lbm_hip.h).  Julia names map as follows: ``collide!``
-> :func:`collide_`, ``stream!`` -> :func:`stream_`, ``apply!`` ->
:func:`apply_`, ``simulate`` -> :func:`simulate`; population arrays are
:class:`HipLattice` (device memory, the reference's (nx, ny, Q) layout).
"""
from ._lib import LbmError, lib  # noqa: F401
from .adjoint import AdjointProblem, divergence_free_field, rrule, simulate_loss  # noqa: F401
from .batch import LatticeBatch  # noqa: F401
from .boundary_conditions import BounceBack, East, MovingWall, North, South, West  # noqa: F401
from .collision_models import (MRT, SRT, TRT, TRT_Λ, IterativeInitializationCollisionModel,  # noqa: F401
                               collision_model_from_viscosity, trt_magic)
from .initial_conditions import (AnalyticalEquilibrium,  # noqa: F401
                                 AnalyticalEquilibriumAndOffEquilibrium,
                                 IterativeInitializationMeiEtAl, cell_centres,
                                 equilibrium_populations, initialize, initialize_analytical,
                                 initialize_mei)
from .lattice import HipLattice  # noqa: F401
from .model import (LatticeBoltzmannModel, ProcessingMethod, apply_boundary_conditions_,  # noqa: F401
                    autotune, collide_model_, next_, simulate, simulate_launch_times, simulate_timed,
                    stream_model_)
from .moments import conserved_sums, density, moments, velocity  # noqa: F401
from .problems import TGV, TaylorGreenVortex  # noqa: F401
from .processing import (DensityConvergence, ExpectedFields, MeanVelocityStoppingCriteria,  # noqa: F401
                         NoStoppingCriteria, StopCriteria, TakeSnapshots, TrackHydrodynamicErrors,
                         TrackRow,
                         VelocityConvergenceStoppingCriteria, VelocityTracker,
                         stop_criteria_for, track_errors, velocity_stats)
from .operators import (apply_, average_clock_mhz, clock_sample, collide_, stream_,  # noqa: F401
                        stream_bc_, synchronize)
from .quadratures import (D2Q4, D2Q5, D2Q9, D2Q13, D2Q17, D2Q21, D2Q37, Quadrature,  # noqa: F401
                          Quadratures, dimension, opposite, order)


def device_count() -> int:
    return int(lib().lbm_device_count())
