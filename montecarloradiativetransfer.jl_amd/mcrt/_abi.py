"""ctypes view of include/mcrt.h and the loader for libmcrt_hip.so.

The structs mirror the C declarations field for field; ``tests/test_abi.py``
checks their sizes and offsets against the C compiler's.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_DIR / "libmcrt_hip.so"
HEADER_PATH = PKG_DIR.parent / "include" / "mcrt.h"

MCRT_OK = 0
MCRT_ERR_INVALID = -1
MCRT_ERR_UNSUPPORTED = -2
MCRT_ERR_NOMEM = -3
MCRT_ERR_HIP = -4
MCRT_ERR_CANCELLED = -5
MCRT_ERR_NOT_READY = -6
MCRT_ERR_NO_DEVICE = -7

SCATTERER_RAYLEIGH = 0
SCATTERER_MIE = 1
SCATTERER_SPLINE = 2
GEOMETRY_PLANE = 0
GEOMETRY_SPHERE = 1

PRECISION_FP64_PLAIN = 0
PRECISION_FP64 = 1
PRECISION_MIXED = 2
PRECISIONS = {"fp64_plain": PRECISION_FP64_PLAIN, "fp64": PRECISION_FP64, "mixed": PRECISION_MIXED}

_dp = C.POINTER(C.c_double)


class mcrt_scatterer(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("n_theta", C.c_int32),
        ("random_points", C.c_int32),
        ("reserved", C.c_int32),
        ("omega", C.c_double),
        ("kl", C.c_double),
        ("csrn", _dp),
        ("F11", _dp),
        ("F12", _dp),
        ("F33", _dp),
        ("F34", _dp),
        ("S11", _dp),
        ("S22", _dp),
        ("F22", _dp),
        ("F44", _dp),
    ]


class mcrt_layer(C.Structure):
    _fields_ = [
        ("d", C.c_double),
        ("n", C.c_double),
        ("tau", C.c_double),
        ("omega", C.c_double),
        ("scatterer", C.POINTER(mcrt_scatterer)),
    ]


class mcrt_geometry(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("n_layers", C.c_int32),
        ("theta", C.c_double),
        ("phi", C.c_double),
        ("layers", C.POINTER(mcrt_layer)),
        ("R", C.c_double),
        ("n", C.c_double),
        ("tau", C.c_double),
        ("omega", C.c_double),
        ("scatterer", C.POINTER(mcrt_scatterer)),
    ]


class mcrt_config(C.Structure):
    _fields_ = [
        ("number_of_rays", C.c_int64),
        ("minimum_intensity", C.c_double),
        ("maximum_scattering_times", C.c_int64),
        ("maximum_threads", C.c_int64),
        ("compute_cb", C.c_int32),
        ("reserved", C.c_int32),
        ("tau0", C.c_double),
        ("n_phi", C.c_int64),
        ("phi", _dp),
        ("cos_phi", _dp),
        ("sin_phi", _dp),
        ("n_theta", C.c_int64),
        ("cos_theta", _dp),
        ("sin_theta", _dp),
        ("inorm", _dp),
        ("n_theta_b", C.c_int64),
        ("theta_b", _dp),
        ("cos_theta_b", _dp),
        ("sin_theta_b", _dp),
        ("n_phi_b", C.c_int64),
        ("phi_b", _dp),
        ("cos_phi_b", _dp),
        ("sin_phi_b", _dp),
    ]


class mcrt_stats(C.Structure):
    _fields_ = [
        ("rays", C.c_uint64),
        ("events", C.c_uint64),
        ("rt_evals", C.c_uint64),
        ("cb_evals", C.c_uint64),
        ("cb_evals_single", C.c_uint64),
        ("rejections", C.c_uint64),
        ("nonfinite", C.c_uint64),
        ("kernel_ms", C.c_double),
        ("total_ms", C.c_double),
        ("batches", C.c_uint64),
        ("chunks", C.c_uint64),
        ("queue_ms", C.c_double),
        ("wait_ms", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class mcrt_request(C.Structure):
    _fields_ = [
        ("cfg", C.POINTER(mcrt_config)),
        ("geom", C.POINTER(mcrt_geometry)),
        ("I0", C.c_double * 4),
        ("seed", C.c_uint64),
        ("ray_begin", C.c_uint64),
        ("instance", C.c_uint32),
        ("precision", C.c_int32),
        ("chunk_rays", C.c_int64),
        ("slots", C.c_int32),
        ("stage_s", C.c_int32),
    ]


class mcrt_launch_shape(C.Structure):
    _fields_ = [("slots", C.c_int32), ("workgroups_per_cu", C.c_int32), ("stage_s", C.c_int32),
                ("reserved", C.c_int32)]


class mcrt_server_stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("submitted", "completed", "cancelled", "failed",
                                          "launches", "chunks", "rays", "queued", "running")] + [
        ("device_bytes_in_use", C.c_int64),
        ("device_budget", C.c_int64),
        ("table_cache_bytes", C.c_int64),
        ("table_hits", C.c_uint64),
        ("table_misses", C.c_uint64),
        ("busy_ms", C.c_double),
        ("preemptions", C.c_uint64),
        ("resumed", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


# mcrt_plan_eval functions (include/mcrt.h MCRT_EVAL_*)
EVAL = {"update_I": 0, "update_E": 1, "update_E_path": 2, "distance": 3, "phase": 4,
        "amplitude": 5, "sample": 6, "phase_transport": 7, "kepler": 8, "phase_diag": 9,
        "bin_math": 10, "cb_phase": 11}

TICKET_STATES = {0: "queued", 1: "running", 2: "done", 3: "cancelled", 4: "failed"}

# Every function declared in include/mcrt.h: name -> (restype, argtypes).
_vp = C.c_void_p
SIGNATURES = {
    "mcrt_version": (C.c_char_p, []),
    "mcrt_last_error": (C.c_char_p, [_vp]),
    "mcrt_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "mcrt_context_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "mcrt_context_destroy": (None, [_vp]),
    "mcrt_table_upload": (C.c_int, [_vp, C.POINTER(mcrt_scatterer), C.POINTER(C.c_uint64)]),
    "mcrt_table_release": (C.c_int, [_vp, C.c_uint64]),
    "mcrt_table_cache_stats": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_uint64)]),
    "mcrt_bhmie": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                             _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "mcrt_calculate_csrn": (C.c_int, [_dp, C.c_int32, C.c_int32, _dp]),
    "mcrt_gauss_legendre": (C.c_int, [C.c_int32, _dp, _dp]),
    "mcrt_initial_weights": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, _dp, _dp, _dp]),
    "mcrt_simulate": (C.c_int, [_vp, C.POINTER(mcrt_config), C.POINTER(mcrt_geometry), _dp,
                                C.c_uint64, C.c_int32, _dp, _dp, _dp, C.POINTER(mcrt_stats)]),
    "mcrt_plan_create": (C.c_int, [_vp, C.POINTER(mcrt_config), C.POINTER(mcrt_geometry), _dp,
                                   C.c_uint64, C.c_int32, C.POINTER(_vp)]),
    "mcrt_plan_launch": (C.c_int, [_vp]),
    "mcrt_plan_wait": (C.c_int, [_vp]),
    "mcrt_plan_read": (C.c_int, [_vp, _dp, _dp, _dp, C.POINTER(mcrt_stats)]),
    "mcrt_plan_kernel_ms": (C.c_int, [_vp, _dp]),
    "mcrt_plan_create_range": (C.c_int, [_vp, C.POINTER(mcrt_config), C.POINTER(mcrt_geometry), _dp,
                                         C.c_uint64, C.c_int32, C.c_uint64, C.c_uint32,
                                         C.POINTER(_vp)]),
    "mcrt_plan_read_sigma": (C.c_int, [_vp, _dp, _dp, _dp, C.POINTER(C.c_int32)]),
    "mcrt_plan_create_shaped": (C.c_int, [_vp, C.POINTER(mcrt_config), C.POINTER(mcrt_geometry), _dp,
                                          C.c_uint64, C.c_int32, C.c_uint64, C.c_uint32,
                                          C.POINTER(mcrt_launch_shape), C.POINTER(_vp)]),
    "mcrt_plan_shape": (C.c_int, [_vp, C.POINTER(mcrt_launch_shape), C.POINTER(C.c_int32)]),
    "mcrt_plan_eval": (C.c_int, [_vp, C.c_int32, C.c_int64, _dp, _dp]),
    "mcrt_simulate_static": (C.c_int, [_vp, C.POINTER(mcrt_config), C.POINTER(mcrt_geometry), _dp,
                                       C.c_uint64, C.c_int32, _dp, _dp, _dp, C.POINTER(mcrt_stats)]),
    "mcrt_batch_create": (C.c_int, [_vp, C.c_int32, C.POINTER(mcrt_request), C.POINTER(_vp)]),
    "mcrt_batch_launch": (C.c_int, [_vp]),
    "mcrt_batch_wait": (C.c_int, [_vp]),
    "mcrt_batch_read": (C.c_int, [_vp, C.c_int32, _dp, _dp, _dp, _dp, _dp, _dp,
                                  C.POINTER(mcrt_stats)]),
    "mcrt_batch_kernel_ms": (C.c_int, [_vp, _dp]),
    "mcrt_batch_info": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mcrt_batch_destroy": (None, [_vp]),
    "mcrt_server_create": (C.c_int, [_vp, C.c_int64, C.c_int64, C.POINTER(_vp)]),
    "mcrt_submit": (C.c_int, [_vp, C.POINTER(mcrt_request), C.POINTER(C.c_uint64)]),
    "mcrt_poll": (C.c_int, [_vp, C.c_uint64, C.POINTER(C.c_int32)]),
    "mcrt_wait": (C.c_int, [_vp, C.c_uint64, C.c_double, C.POINTER(C.c_int32)]),
    "mcrt_result": (C.c_int, [_vp, C.c_uint64, _dp, _dp, _dp, _dp, _dp, _dp,
                              C.POINTER(mcrt_stats), _dp]),
    "mcrt_cancel": (C.c_int, [_vp, C.c_uint64]),
    "mcrt_server_stats_get": (C.c_int, [_vp, C.POINTER(mcrt_server_stats)]),
    "mcrt_server_destroy": (None, [_vp]),
    "mcrt_server_set_slice": (C.c_int, [_vp, C.c_double]),
    "mcrt_plan_destroy": (None, [_vp]),
}

_LIB = None


class MCRTError(RuntimeError):
    """A non-zero status from libmcrt_hip.so."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class ArgumentError(MCRTError, ValueError):
    """MCRT_ERR_INVALID: the reference raises ArgumentError for these."""


def load_library(path: os.PathLike | None = None):
    """Load libmcrt_hip.so (once).  Raises when it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the transport has no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def check(code: int, ctx=None) -> None:
    if code == MCRT_OK:
        return
    lib = load_library()
    msg = lib.mcrt_last_error(ctx).decode("utf-8", "replace")
    if code == MCRT_ERR_INVALID:
        raise ArgumentError(code, msg)
    raise MCRTError(code, msg)


def dptr(a: np.ndarray | None):
    """A double* into a C-contiguous float64 array (NULL for None)."""
    if a is None:
        return _dp()
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "expected contiguous float64"
    return a.ctypes.data_as(_dp)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))
