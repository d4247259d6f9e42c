"""ctypes binding of ``include/rayfactor.h`` (the library's C ABI).

The native library must be built (``python -m rayfactorcl_amd.build`` or
``__graft_entry__.build()``); importing this module without it raises, there
is no Python fallback.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_float, c_int32, c_int64, c_size_t, c_uint32, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().with_name("librayfactor_amd.so")

# Byte layouts of RayfactorConstants.h:38-56 (112 and 72 bytes).
PRIM_DTYPE = np.dtype([
    ("m", "<f4", (12,)), ("invm", "<f4", (12,)), ("type", "<i4"), ("rayDensity", "<i4"),
    ("util1", "<f4"), ("util2", "<f4"),
])
BV_DTYPE = np.dtype([
    ("invm", "<f4", (12,)), ("type", "<i4"), ("settings", "<i4"), ("objStartIndex", "<i4"),
    ("objEndIndex", "<i4"), ("util1", "<f4"), ("util2", "<f4"),
])
assert PRIM_DTYPE.itemsize == 112 and BV_DTYPE.itemsize == 72

CYLINDER, FRUSTUM, SPHERE, ANNULUS, RECTANGLE, DISC, NONE = range(7)

RF_OK = 0
RF_ERR_INVALID_ARG = -1
RF_ERR_INVALID_SCENE = -2
RF_ERR_HIP = -3
RF_ERR_OUT_OF_MEMORY = -4
RF_ERR_OVERFLOW = -5
RF_ERR_IO = -6
RF_ERR_PARSE = -7
RF_ERR_CAPACITY = -8
RF_ERR_NO_DEVICE = -9

# rf_timing.shape_flags
RF_SHAPE_MFMA = 1
RF_SHAPE_STACKS = 2
RF_SHAPE_SEARCHED = 4
RF_SHAPE_BAND_TILE = 8
RF_SHAPE_COUNTED = 16
RF_SHAPE_BV_BINNED = 32
RF_SHAPE_REFERENCE = 64
RF_SHAPE_HIERARCHY = 128

MATH = {"compat": 0, "fp64": 1, "fast": 2}
RNG = {"reference": 0, "seeded": 1}
FORM = {"per_emitter": 0, "persistent": 1, "reference": 2}
TRANSFORM = {"auto": 0, "valu": 1, "mfma": 2}


class SolveOpts(ctypes.Structure):
    _fields_ = [
        ("math", c_int32), ("rng", c_int32), ("use_bvs", c_int32), ("budget", c_int32),
        ("n_samples", c_int32), ("form", c_int32), ("rays_per_emitter", c_int64),
        ("seed", c_uint32), ("geometry", c_uint32), ("refine", c_int32),
        ("rays_per_thread", c_int32), ("block_size", c_int32), ("transform", c_int32),
        ("refine_margin", c_float), ("per_test", c_int32), ("count_tests", c_int32), ("band_tile", c_int32),
        ("hierarchy", c_int32),
    ]


class Timing(ctypes.Structure):
    _fields_ = [
        ("kernel_ms", c_double), ("wall_ms", c_double), ("rays", c_int64), ("tests", c_int64),
        ("flops", c_int64), ("refined", c_int64), ("launches", c_int64), ("rescans", c_int64),
        ("graph_launches", c_int64), ("tests_brute_force", c_int64), ("shape_rays_per_thread", c_int32),
        ("shape_block", c_int32), ("shape_flags", c_int32), ("shape_pad", c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class HostScene(ctypes.Structure):
    _fields_ = [
        ("prims", c_void_p), ("n_prims", c_int32), ("bvs", c_void_p), ("n_bvs", c_int32),
        ("globalRayDensity", c_float), ("useBoundingVolumes", c_int32),
    ]


class VfReport(ctypes.Structure):
    _fields_ = [
        ("min_closure", c_double), ("max_closure", c_double), ("max_recip_abs", c_double),
        ("max_recip_rel", c_double), ("worst_i", c_int32), ("worst_j", c_int32), ("pairs_compared", c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class RayFactorError(RuntimeError):
    def __init__(self, code: int, where: str, detail: str = ""):
        self.code = code
        super().__init__(f"{where}: status {code} ({status_string(code)}){': ' + detail if detail else ''}")


# Every function declared in include/rayfactor.h: name -> (restype, argtypes).
_P = c_void_p
SIGNATURES = {
    "rf_abi_version": (c_int32, []),
    "rf_status_string": (c_char_p, [c_int32]),
    "rf_solve_opts_default": (None, [POINTER(SolveOpts)]),
    "rf_ctx_create": (c_int32, [c_int32, POINTER(_P)]),
    "rf_ctx_destroy": (c_int32, [_P]),
    "rf_last_error": (c_char_p, [_P]),
    "rf_device_info": (c_int32, [_P, c_char_p, c_size_t, POINTER(c_int32), POINTER(c_int32)]),
    "rf_scene_create": (c_int32, [_P, _P, c_int32, _P, c_int32, POINTER(_P)]),
    "rf_scene_destroy": (c_int32, [_P]),
    "rf_scene_size": (c_int32, [_P, POINTER(c_int32), POINTER(c_int32)]),
    "rf_scene_footprint": (c_int64, [c_int32, c_int32]),
    "rf_nthreads": (c_int32, [c_int32, c_int32, c_int32, c_int32, POINTER(c_int64)]),
    "rf_launch_emitter": (c_int32, [_P, _P, c_int32, c_int64, c_int64, POINTER(SolveOpts)]),
    "rf_sync": (c_int32, [_P]),
    "rf_hits_read": (c_int32, [_P, _P, _P]),
    "rf_vf_read": (c_int32, [_P, _P, c_int64, _P]),
    "rf_hits_reset": (c_int32, [_P, _P]),
    "rf_solve": (c_int32, [_P, _P, POINTER(SolveOpts), _P, _P, POINTER(Timing)]),
    "rf_solve_noalloc": (c_int32, [_P, _P, POINTER(SolveOpts), _P]),
    "rf_trace_rays": (c_int32, [_P, _P, c_int32, c_int64, c_int64, POINTER(SolveOpts), _P]),
    "rf_run_cl_gpurng": (c_int32, [_P, _P, c_int32, _P, c_int32, c_int32, c_int32, _P, _P, _P]),
    "rf_batch_solve": (c_int32, [_P, _P, c_int32, POINTER(SolveOpts), _P, _P, POINTER(Timing)]),
    "rf_autotune": (c_int32, [_P, _P, c_int32, POINTER(SolveOpts), POINTER(c_int32), POINTER(c_int32)]),
    "rf_stream_create": (c_int32, [_P, c_int64, POINTER(_P)]),
    "rf_stream_destroy": (c_int32, [_P]),
    "rf_stream_submit": (c_int32, [_P, _P, c_int32, POINTER(SolveOpts), POINTER(c_int64)]),
    "rf_stream_flush": (c_int32, [_P, POINTER(c_int32)]),
    "rf_stream_step": (c_int32, [_P, POINTER(c_int32)]),
    "rf_stream_collect": (c_int32, [_P, c_int64, _P, _P, POINTER(c_double)]),
    "rf_stream_resident_bytes": (c_int64, [_P]),
    "rf_stream_set_cost_limits": (c_int32, [_P, c_double, c_double]),
    "rf_stream_queued_cost": (c_double, [_P]),
    "rf_opts_same_tallies": (c_int32, [POINTER(SolveOpts), POINTER(SolveOpts)]),
    "rf_debug_fail_solve": (c_int32, [_P, c_int32]),
    "rf_scene_load_xml": (c_int32, [c_char_p, POINTER(HostScene)]),
    "rf_scene_parse_xml": (c_int32, [c_char_p, c_size_t, POINTER(HostScene)]),
    "rf_host_scene_free": (None, [POINTER(HostScene)]),
    "rf_write_output": (c_int32, [c_char_p, c_int32, _P, _P]),
    "rf_normalise": (c_int32, [c_int32, _P, _P, _P]),
    "rf_selftest_math": (c_int32, [_P, _P, _P]),
    "rf_selftest_sincos": (c_int32, [_P, _P]),
    "rf_areas": (c_int32, [_P, c_int32, _P]),
    "rf_vf_check": (c_int32, [c_int32, _P, _P, c_double, _P, POINTER(VfReport)]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """The loaded native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m rayfactorcl_amd.build` "
                "(there is no fallback implementation)")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def status_string(code: int) -> str:
    try:
        return lib().rf_status_string(code).decode()
    except Exception:  # library unavailable
        return "unknown"


def check(code: int, where: str, ctx=None) -> None:
    if code != RF_OK:
        detail = ""
        try:
            detail = (lib().rf_last_error(ctx) or b"").decode()
        except Exception:
            pass
        raise RayFactorError(code, where, detail)


def ptr(a: np.ndarray | None) -> int | None:
    return None if a is None else a.ctypes.data
