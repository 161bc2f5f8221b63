"""ctypes binding of liblbm_hip.so (the C ABI in include/lbm_hip.h).

The library is the only compute path: if it is missing this module raises
instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("LBM_HIP_LIB",
                          os.path.join(os.path.dirname(_HERE), "lib", "liblbm_hip.so"))

STORE_F64, STORE_F32, STORE_F16DEV, STORE_BF16DEV, STORE_F16NEQ, STORE_BF16NEQ = 0, 1, 2, 3, 4, 5
STORAGES = {"f64": STORE_F64, "f32": STORE_F32, "f16dev": STORE_F16DEV, "bf16dev": STORE_BF16DEV,
            "f16neq": STORE_F16NEQ, "bf16neq": STORE_BF16NEQ}
SRT_KIND, TRT_KIND, MRT_KIND, MEI_KIND = 0, 1, 2, 3
FORCE_NONE, FORCE_CONST, FORCE_FIELD = 0, 1, 2
BOUNCE_BACK, MOVING_WALL = 0, 1
NORTH, EAST, SOUTH, WEST = 0, 1, 2, 3

# every entry point declared in include/lbm_hip.h
EXPORTS = [
    "lbm_last_error", "lbm_version", "lbm_device_count", "lbm_synchronize",
    "lbm_clock_sample", "lbm_quadrature_info", "lbm_quadrature_f32",
    "lbm_field_create", "lbm_field_destroy", "lbm_field_info", "lbm_field_upload",
    "lbm_field_download", "lbm_field_copy", "lbm_initialize_analytical",
    "lbm_model_create", "lbm_model_destroy", "lbm_model_set_srt", "lbm_model_set_trt",
    "lbm_model_set_mrt", "lbm_model_set_force", "lbm_model_set_bcs", "lbm_model_set_launch",
    "lbm_autotune",
    "lbm_collide", "lbm_stream", "lbm_apply_bcs", "lbm_stream_bc", "lbm_simulate",
    "lbm_simulate_timed", "lbm_simulate_launch_times", "lbm_moments", "lbm_reduce",
    "lbm_reduce_noalloc",
    "lbm_batch_create", "lbm_batch_destroy", "lbm_batch_info", "lbm_batch_upload",
    "lbm_batch_download", "lbm_batch_set_launch", "lbm_batch_set_nodes_per_thread",
    "lbm_batch_get_launch", "lbm_batch_autotune", "lbm_batch_simulate",
    "lbm_snapshots_create", "lbm_snapshots_destroy", "lbm_snapshot_take", "lbm_snapshots_wait",
    "lbm_snapshot_get",
    "lbm_adjoint_create", "lbm_adjoint_destroy", "lbm_adjoint_info", "lbm_adjoint_loss",
    "lbm_adjoint_gradient",
    "lbm_expected_create", "lbm_expected_destroy", "lbm_expected_upload", "lbm_track_errors",
    "lbm_velocity_tracker_create", "lbm_velocity_tracker_destroy", "lbm_velocity_stats",
    "lbm_model_set_mei", "lbm_field_node", "lbm_mei_initialize",
]


class LbmBC(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("direction", ctypes.c_int32),
        ("x0", ctypes.c_int32),
        ("x1", ctypes.c_int32),
        ("y0", ctypes.c_int32),
        ("y1", ctypes.c_int32),
        ("ux", ctypes.c_double),
        ("uy", ctypes.c_double),
        ("rho", ctypes.c_double),
    ]


class LbmTrackParams(ctypes.Structure):
    _fields_ = [
        ("tau", ctypes.c_double),
        ("u_max", ctypes.c_double),
        ("delta", ctypes.c_double),
        ("affine", ctypes.c_double * 16),
    ]


class LbmError(RuntimeError):
    """An lbm_* call returned a negative status."""


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python latticeboltzmann.jl_amd/build.py` "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    i = ctypes.c_int
    d = ctypes.c_double
    i64 = ctypes.c_int64
    L.lbm_last_error.restype = ctypes.c_char_p
    L.lbm_last_error.argtypes = []
    L.lbm_version.restype = i
    L.lbm_device_count.restype = i
    L.lbm_synchronize.restype = i
    L.lbm_clock_sample.argtypes = [ctypes.POINTER(ctypes.c_uint64), dp]
    L.lbm_quadrature_info.argtypes = [i, ip, dp, dp, dp, dp, ip, ip]
    L.lbm_quadrature_f32.argtypes = [i, vp, vp]
    L.lbm_field_create.argtypes = [ctypes.POINTER(vp), i, i, i, i, d]
    L.lbm_field_destroy.argtypes = [vp]
    L.lbm_field_info.argtypes = [vp, ip, ip, ip, ip, ip]
    L.lbm_field_upload.argtypes = [vp, dp]
    L.lbm_field_download.argtypes = [vp, dp]
    L.lbm_field_copy.argtypes = [vp, vp]
    L.lbm_initialize_analytical.argtypes = [vp, i, dp, dp, dp, dp, dp, d, d]
    L.lbm_model_create.argtypes = [ctypes.POINTER(vp), i]
    L.lbm_model_destroy.argtypes = [vp]
    L.lbm_model_set_srt.argtypes = [vp, d]
    L.lbm_model_set_trt.argtypes = [vp, d, d]
    L.lbm_model_set_mrt.argtypes = [vp, dp, i]
    L.lbm_model_set_force.argtypes = [vp, i, dp, i, i]
    L.lbm_model_set_bcs.argtypes = [vp, ctypes.POINTER(LbmBC), i]
    L.lbm_model_set_launch.argtypes = [vp, i, i, i, i]
    L.lbm_autotune.argtypes = [vp, vp, i, ip, ip, ip, ip, dp]
    for name in ["lbm_collide", "lbm_stream", "lbm_apply_bcs", "lbm_stream_bc"]:
        getattr(L, name).argtypes = [vp, vp, vp]
    L.lbm_simulate.argtypes = [vp, vp, vp, i64]
    L.lbm_simulate_timed.argtypes = [vp, vp, vp, i64, dp, ctypes.POINTER(i64)]
    L.lbm_simulate_launch_times.argtypes = [vp, vp, vp, i64, dp, ctypes.POINTER(i64)]
    L.lbm_moments.argtypes = [vp, d, dp, dp, dp, dp, dp, dp]
    L.lbm_reduce.argtypes = [vp, dp]
    L.lbm_reduce_noalloc.argtypes = [vp, dp]
    L.lbm_snapshots_create.argtypes = [ctypes.POINTER(vp), i, i, i, i, ctypes.c_char_p]
    L.lbm_snapshots_destroy.argtypes = [vp]
    L.lbm_snapshot_take.argtypes = [vp, vp, i64]
    L.lbm_snapshots_wait.argtypes = [vp, ip]
    L.lbm_snapshot_get.argtypes = [vp, i, ctypes.POINTER(i64), dp]
    L.lbm_batch_create.argtypes = [ctypes.POINTER(vp), i, i, i, ip, ip, dp, dp, d]
    L.lbm_batch_destroy.argtypes = [vp]
    L.lbm_batch_info.argtypes = [vp, ip, ctypes.POINTER(i64), ctypes.POINTER(i64)]
    L.lbm_batch_upload.argtypes = [vp, i, dp]
    L.lbm_batch_download.argtypes = [vp, i, i, dp]
    L.lbm_batch_set_launch.argtypes = [vp, i, i]
    L.lbm_batch_set_nodes_per_thread.argtypes = [vp, i]
    L.lbm_batch_get_launch.argtypes = [vp, ip, ip, ip]
    L.lbm_batch_autotune.argtypes = [vp, i, ip, ip, dp]
    L.lbm_batch_simulate.argtypes = [vp, i64, i, dp]
    L.lbm_adjoint_create.argtypes = [ctypes.POINTER(vp), vp, i, i, i, d, i64, i]
    L.lbm_adjoint_destroy.argtypes = [vp]
    L.lbm_adjoint_info.argtypes = [vp, ip, ip, ctypes.POINTER(i64)]
    L.lbm_adjoint_loss.argtypes = [vp, dp, dp, dp, dp]
    L.lbm_adjoint_gradient.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.lbm_model_set_mei.argtypes = [vp, d, dp, i, i]
    L.lbm_field_node.argtypes = [vp, i, i, dp]
    L.lbm_mei_initialize.argtypes = [vp, vp, vp, d, i64, ctypes.POINTER(i64)]
    L.lbm_expected_create.argtypes = [ctypes.POINTER(vp), i, i]
    L.lbm_expected_destroy.argtypes = [vp]
    L.lbm_expected_upload.argtypes = [vp, dp]
    L.lbm_track_errors.argtypes = [vp, vp, ctypes.POINTER(LbmTrackParams), dp]
    L.lbm_velocity_tracker_create.argtypes = [ctypes.POINTER(vp), i, i]
    L.lbm_velocity_tracker_destroy.argtypes = [vp]
    L.lbm_velocity_stats.argtypes = [vp, vp, dp]
    for name in EXPORTS:
        f = getattr(L, name)
        if name not in ("lbm_last_error",):
            f.restype = ctypes.c_int
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc < 0:
        msg = lib().lbm_last_error().decode(errors="replace")
        raise LbmError(f"{what}: {msg} (status {rc})" if what else f"{msg} (status {rc})")


def dptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
