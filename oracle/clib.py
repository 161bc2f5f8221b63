"""ctypes wrapper around ``oracle/liblbm_oracle.so`` (the C restatement).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Arrays follow the reference's layout: ``numpy`` float64 arrays of shape
``(nx, ny, Q)`` in Fortran (column-major) order, i.e. a Julia
``Array{Float64,3}`` (src/initial_conditions.jl:13).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblbm_oracle.so")

QMAX = 37
FORCE_NONE, FORCE_CONST, FORCE_FIELD = 0, 1, 2
BOUNCE_BACK, MOVING_WALL = 0, 1
NORTH, EAST, SOUTH, WEST = 0, 1, 2, 3
SRT, TRT, MRT = 0, 1, 2


class Quad(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_int),
        ("Q", ctypes.c_int),
        ("cx", ctypes.c_int * QMAX),
        ("cy", ctypes.c_int * QMAX),
        ("w", ctypes.c_double * QMAX),
        ("s", ctypes.c_double),
        ("order", ctypes.c_int),
        ("feq_order", ctypes.c_int),
        ("opp", ctypes.c_int * QMAX),
    ]


class BC(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("direction", ctypes.c_int),
        ("x0", ctypes.c_int),
        ("x1", ctypes.c_int),
        ("y0", ctypes.c_int),
        ("y1", ctypes.c_int),
        ("ux", ctypes.c_double),
        ("uy", ctypes.c_double),
        ("rho", ctypes.c_double),
    ]


_lib = None


def build() -> None:
    """Compile the C restatement with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        qp = ctypes.POINTER(Quad)
        bp = ctypes.POINTER(BC)
        _lib.lbo_quadrature.argtypes = [ctypes.c_int, qp]
        _lib.lbo_density.argtypes = [qp, dp]
        _lib.lbo_density.restype = ctypes.c_double
        _lib.lbo_velocity.argtypes = [qp, dp, ctypes.c_double, dp]
        _lib.lbo_equilibrium.argtypes = [qp, ctypes.c_double, dp, dp]
        _lib.lbo_hermite_equilibrium.argtypes = [qp, ctypes.c_double, dp, dp]
        _lib.lbo_init_node.argtypes = [qp, ctypes.c_int, ctypes.c_double, dp, ctypes.c_double, dp,
                                       ctypes.c_double, ctypes.c_double, dp]
        _lib.lbo_initialize.argtypes = [qp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp,
                                        dp, dp, ctypes.c_double, ctypes.c_double, dp]
        i = ctypes.c_int
        d = ctypes.c_double
        _lib.lbo_collide_srt.argtypes = [qp, i, i, dp, dp, d, i, dp]
        _lib.lbo_collide_trt.argtypes = [qp, i, i, dp, dp, d, d, i, dp]
        _lib.lbo_collide_mrt.argtypes = [qp, i, i, dp, dp, dp, i, i, dp]
        _lib.lbo_stream.argtypes = [qp, i, i, dp, dp]
        _lib.lbo_apply_bcs.argtypes = [qp, i, i, bp, i, dp, dp]
        _lib.lbo_apply_bcs.restype = i
        _lib.lbo_simulate.restype = i
        _lib.lbo_simulate.argtypes = [qp, i, i, dp, dp, i, dp, i, i, dp, bp, i, ctypes.c_long]
        _lib.lbo_moments.argtypes = [qp, i, i, dp, d, dp, dp, dp, dp, dp, dp]
        _lib.lbo_track_errors.argtypes = [qp, i, i, dp, dp, dp, d, d, d, dp]
        _lib.lbo_velocity_stats.argtypes = [qp, i, i, dp, dp, dp]
        _lib.lbo_mei_nonlinear.argtypes = [qp, i, i, dp, dp]
        _lib.lbo_collide_mei.argtypes = [qp, i, i, dp, dp, d, dp]
        _lib.lbo_mei_initialize.argtypes = [qp, i, i, dp, d, d, ctypes.c_long, bp, i, dp, dp]
        _lib.lbo_mei_initialize.restype = ctypes.c_long
        _lib.lbo_set_threads.argtypes = [i]
        _lib.lbo_get_threads.restype = i
    return _lib


def _dp(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def quadrature(qid: int) -> Quad:
    q = Quad()
    if lib().lbo_quadrature(qid, ctypes.byref(q)) != 0:
        raise ValueError(f"unknown quadrature D2Q{qid}")
    return q


def quad_arrays(qid: int):
    """(c [2,Q] int, w [Q] float64, s, order, opp [Q] 0-based)"""
    q = quadrature(qid)
    Q = q.Q
    c = np.array([list(q.cx)[:Q], list(q.cy)[:Q]], dtype=np.int64)
    w = np.array(list(q.w)[:Q])
    return c, w, q.s, q.order, np.array(list(q.opp)[:Q]), q.feq_order


def set_threads(n: int) -> None:
    lib().lbo_set_threads(int(n))


def _check_f(f, Q=None):
    if not (isinstance(f, np.ndarray) and f.dtype == np.float64 and f.ndim == 3
            and f.flags.f_contiguous):
        raise TypeError("populations must be float64 (nx, ny, Q) Fortran-ordered arrays")
    if Q is not None and f.shape[2] != Q:
        raise ValueError("population count does not match the quadrature")


def _force_args(force, nx, ny):
    if force is None:
        return FORCE_NONE, None, None
    F = np.asarray(force, dtype=np.float64)
    if F.size == 2:
        keep = np.ascontiguousarray(F.reshape(2))
        return FORCE_CONST, _dp(keep), keep
    keep = np.asfortranarray(F.reshape(nx, ny, 2))
    return FORCE_FIELD, _dp(keep), keep


def equilibrium(qid, rho, u):
    q = quadrature(qid)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(q.Q)
    lib().lbo_equilibrium(ctypes.byref(q), float(rho), _dp(uu), _dp(out))
    return out


def hermite_equilibrium(qid, rho, u):
    q = quadrature(qid)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(q.Q)
    lib().lbo_hermite_equilibrium(ctypes.byref(q), float(rho), _dp(uu), _dp(out))
    return out


def init_node(qid, mode, rho, u, T=1.0, G=None, tau=0.0, factor=1.0):
    """One node of the analytic initial conditions (lbo_init_node)."""
    q = quadrature(qid)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    g = np.ascontiguousarray(np.zeros(4) if G is None else np.asarray(G, dtype=np.float64)
                             .reshape(2, 2).ravel(order="F"))
    out = np.zeros(q.Q)
    lib().lbo_init_node(ctypes.byref(q), int(mode), float(rho), _dp(uu), float(T), _dp(g),
                        float(tau), float(factor), _dp(out))
    return out


def initialize(qid, mode, rho, ux, uy, T=None, G=None, tau=0.0, factor=1.0):
    """initialize(strategy, q, problem) for the analytic strategies over a
    lattice: planes (nx, ny); G (nx, ny, 2, 2) with G[..., a, b]; returns
    (nx, ny, Q) column-major."""
    q = quadrature(qid)
    rho = np.asfortranarray(rho, dtype=np.float64)
    nx, ny = rho.shape
    ux = np.asfortranarray(ux, dtype=np.float64)
    uy = np.asfortranarray(uy, dtype=np.float64)
    Tp = None if T is None else np.asfortranarray(T, dtype=np.float64)
    Gp = None if G is None else np.asfortranarray(
        np.asarray(G, dtype=np.float64).reshape(nx, ny, 4, order="F"))
    out = np.zeros((nx, ny, q.Q), order="F")
    lib().lbo_initialize(ctypes.byref(q), nx, ny, int(mode), _dp(rho), _dp(ux), _dp(uy),
                         None if Tp is None else _dp(Tp), None if Gp is None else _dp(Gp),
                         float(tau), float(factor), _dp(out))
    return out


def density(qid, f):
    q = quadrature(qid)
    ff = np.ascontiguousarray(f, dtype=np.float64)
    return lib().lbo_density(ctypes.byref(q), _dp(ff))


def velocity(qid, f, rho):
    q = quadrature(qid)
    ff = np.ascontiguousarray(f, dtype=np.float64)
    u = np.zeros(2)
    lib().lbo_velocity(ctypes.byref(q), _dp(ff), float(rho), _dp(u))
    return u


def collide_srt(qid, fin, fout, tau, force=None):
    q = quadrature(qid)
    _check_f(fin, q.Q)
    _check_f(fout, q.Q)
    nx, ny, _ = fin.shape
    k, p, _keep = _force_args(force, nx, ny)
    lib().lbo_collide_srt(ctypes.byref(q), nx, ny, _dp(fin), _dp(fout), float(tau), k, p)


def collide_trt(qid, fin, fout, tau_s, tau_a, force=None):
    q = quadrature(qid)
    _check_f(fin, q.Q)
    _check_f(fout, q.Q)
    nx, ny, _ = fin.shape
    k, p, _keep = _force_args(force, nx, ny)
    lib().lbo_collide_trt(ctypes.byref(q), nx, ny, _dp(fin), _dp(fout), float(tau_s),
                          float(tau_a), k, p)


def collide_mrt(qid, fin, fout, taus, force=None):
    q = quadrature(qid)
    _check_f(fin, q.Q)
    _check_f(fout, q.Q)
    nx, ny, _ = fin.shape
    t = np.ascontiguousarray(taus, dtype=np.float64)
    k, p, _keep = _force_args(force, nx, ny)
    lib().lbo_collide_mrt(ctypes.byref(q), nx, ny, _dp(fin), _dp(fout), _dp(t), t.size, k, p)


def stream(qid, f, fnew):
    q = quadrature(qid)
    _check_f(f, q.Q)
    _check_f(fnew, q.Q)
    nx, ny, _ = f.shape
    lib().lbo_stream(ctypes.byref(q), nx, ny, _dp(f), _dp(fnew))


def make_bcs(bcs):
    """bcs: list of dicts {kind, direction, xs=(x0,x1), ys=(y0,y1), u=(ux,uy), rho}"""
    arr = (BC * max(1, len(bcs)))()
    for k, b in enumerate(bcs):
        arr[k].kind = b["kind"]
        arr[k].direction = b["direction"]
        arr[k].x0, arr[k].x1 = b["xs"]
        arr[k].y0, arr[k].y1 = b["ys"]
        ux, uy = b.get("u", (0.0, 0.0))
        arr[k].ux, arr[k].uy = float(ux), float(uy)
        arr[k].rho = float(b.get("rho", 1.0))
    return arr


def apply_bcs(qid, bcs, fnew, fold):
    q = quadrature(qid)
    _check_f(fnew, q.Q)
    _check_f(fold, q.Q)
    nx, ny, _ = fnew.shape
    arr = make_bcs(bcs)
    if lib().lbo_apply_bcs(ctypes.byref(q), nx, ny, arr, len(bcs), _dp(fnew), _dp(fold)) != 0:
        raise IndexError("boundary range outside the lattice (BoundsError in the reference)")


def simulate(qid, f_stream, f_collision, collision, taus, nsteps, force=None, bcs=()):
    """Run ``nsteps`` iterations of collide -> stream -> BCs in place."""
    q = quadrature(qid)
    _check_f(f_stream, q.Q)
    _check_f(f_collision, q.Q)
    nx, ny, _ = f_stream.shape
    t = np.ascontiguousarray(taus, dtype=np.float64)
    k, p, _keep = _force_args(force, nx, ny)
    arr = make_bcs(list(bcs))
    if lib().lbo_simulate(ctypes.byref(q), nx, ny, _dp(f_stream), _dp(f_collision), collision,
                          _dp(t), t.size, k, p, arr, len(bcs), int(nsteps)) != 0:
        raise IndexError("boundary range outside the lattice (BoundsError in the reference)")


def moments(qid, f, tau):
    """Returns rho, ux, uy, sxx, sxy, syy as (nx, ny) Fortran arrays."""
    q = quadrature(qid)
    _check_f(f, q.Q)
    nx, ny, _ = f.shape
    outs = [np.zeros((nx, ny), order="F") for _ in range(6)]
    lib().lbo_moments(ctypes.byref(q), nx, ny, _dp(f), float(tau), *[_dp(o) for o in outs])
    return outs


def track_errors(qid, f, expected, affine, tau, u_max, delta=1.0):
    """One TrackHydrodynamicErrors row (track_hydrodynamic_errors.jl:52-222):
    error_rho, error_u, error_p, error_sxx, error_sxy, error_syy, error_syx,
    mass, momentum, energy, non-finite nodes.  expected: (nx, ny, 8), planes
    rho, ux, uy, p, sxx, sxy, syx, syy."""
    q = quadrature(qid)
    _check_f(f, q.Q)
    nx, ny, _ = f.shape
    e = np.asfortranarray(expected, dtype=np.float64)
    assert e.shape == (nx, ny, 8)
    a = np.ascontiguousarray(affine, dtype=np.float64)
    assert a.size == 16
    out = np.zeros(11)
    lib().lbo_track_errors(ctypes.byref(q), nx, ny, _dp(f), _dp(e), _dp(a), float(tau),
                           float(u_max), float(delta), _dp(out))
    return out


def velocity_stats(qid, f, u_old=None):
    """Stopping-criteria sums (stopping_criteria.jl:17-99): mean ux, mean uy,
    sum |u - u_old|^2, sum |u_old|^2, non-finite nodes.  u_old: (nx, ny, 2)
    Fortran array, updated in place to u."""
    q = quadrature(qid)
    _check_f(f, q.Q)
    nx, ny, _ = f.shape
    if u_old is not None:
        assert u_old.shape == (nx, ny, 2) and u_old.flags.f_contiguous
        assert u_old.dtype == np.float64
    out = np.zeros(5)
    lib().lbo_velocity_stats(ctypes.byref(q), nx, ny, _dp(f),
                             None if u_old is None else _dp(u_old), _dp(out))
    return out


def mei_nonlinear(qid, u):
    """nonlinear_term of IterativeInitializationCollisionModel
    (iterative_initialization.jl:12-40) for the lattice velocity u (nx, ny, 2)."""
    q = quadrature(qid)
    u = np.asfortranarray(u, dtype=np.float64)
    nx, ny, _ = u.shape
    N = np.zeros((nx, ny, q.Q), order="F")
    lib().lbo_mei_nonlinear(ctypes.byref(q), nx, ny, _dp(u), _dp(N))
    return N


def collide_mei(qid, fin, fout, tau, N):
    q = quadrature(qid)
    _check_f(fin, q.Q)
    _check_f(fout, q.Q)
    nx, ny, _ = fin.shape
    N = np.asfortranarray(N, dtype=np.float64)
    lib().lbo_collide_mei(ctypes.byref(q), nx, ny, _dp(fin), _dp(fout), float(tau), _dp(N))


def mei_initialize(qid, u, tau=1.0, eps=1e-7, max_steps=10000, bcs=()):
    """initialize(IterativeInitializationMeiEtAl(tau, eps), q, problem)
    (mei_et_al.jl:13-41): returns (f_stream, f_collision, steps run)."""
    q = quadrature(qid)
    u = np.asfortranarray(u, dtype=np.float64)
    nx, ny, _ = u.shape
    fs = np.zeros((nx, ny, q.Q), order="F")
    fc = np.zeros((nx, ny, q.Q), order="F")
    arr = make_bcs(list(bcs))
    steps = lib().lbo_mei_initialize(ctypes.byref(q), nx, ny, _dp(u), float(tau), float(eps),
                                     int(max_steps), arr, len(bcs), _dp(fs), _dp(fc))
    if steps < 0:
        raise IndexError("boundary range outside the lattice (BoundsError in the reference)")
    return fs, fc, int(steps)
