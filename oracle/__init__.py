"""CPU oracle for LatticeBoltzmann.jl's collide-stream step.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg may import this package, and only as the
checker.  The product (``liblbm_hip.so`` and the ``lbmhip`` package) never
imports, links or calls anything under ``oracle/``.

Two independent restatements of the reference live here:

* ``liblbm_oracle.so`` (``lbm_oracle.c``): scalar FP64 C that follows the
  reference's evaluation order line by line (wrapped by :mod:`oracle.clib`);
* :mod:`oracle.lbm_numpy`: a vectorised NumPy formulation (``np.roll``
  streaming, boolean-mask boundary conditions, tensor-Hermite equilibria) that
  shares no code with the C file and is used to cross-check it.

Parity status per function is listed in the module docstrings; every function
is pinned by at least one check in ``tests/test_oracle_*.py``.
"""
