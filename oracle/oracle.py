"""ctypes binding of the CPU oracle (oracle/librf_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
check its result and by bench.py's cpu_baseline leg.  The product never uses it.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int32, c_int64, c_uint32, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "librf_oracle.so"

PRIM_DTYPE = np.dtype([
    ("m", "<f4", (12,)), ("invm", "<f4", (12,)), ("type", "<i4"), ("rayDensity", "<i4"),
    ("util1", "<f4"), ("util2", "<f4"),
])
BV_DTYPE = np.dtype([
    ("invm", "<f4", (12,)), ("type", "<i4"), ("settings", "<i4"), ("objStartIndex", "<i4"),
    ("objEndIndex", "<i4"), ("util1", "<f4"), ("util2", "<f4"),
])

MODE = {"compat": 0, "fp64": 1}
RNG = {"reference": 0, "seeded": 1}


class Opts(ctypes.Structure):
    _fields_ = [("mode", c_int32), ("use_bvs", c_int32), ("rng", c_int32), ("seed", c_uint32),
                ("geometry", c_uint32), ("n_threads", c_int32)]


def _load():
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
    L = ctypes.CDLL(str(LIB_PATH))
    P = c_void_p
    sig = {
        "rfo_philox4x32_10": (None, [P, P, P]),
        "rfo_u01_24": (c_float, [c_uint32]),
        "rfo_sincosf": (None, [c_float, POINTER(c_float), POINTER(c_float)]),
        "rfo_sincos64": (None, [c_double, POINTER(c_double), POINTER(c_double)]),
        "rfo_workitem_u32": (None, [c_int32, c_uint32, c_uint32, c_int32, c_int64, P]),
        "rfo_emit_f32": (None, [P, P, P]),
        "rfo_emit_f64": (None, [P, P, P]),
        "rfo_hit_f32": (c_float, [c_int32, P, c_float, c_float, P]),
        "rfo_hit_f64": (c_double, [c_int32, P, c_float, c_float, P]),
        "rfo_find_f32": (c_int32, [P, c_int32, c_int32, P]),
        "rfo_find_f64": (c_int32, [P, c_int32, c_int32, P]),
        "rfo_find_bv_f32": (c_int32, [P, P, c_int32, P]),
        "rfo_find_bv_f64": (c_int32, [P, P, c_int32, P]),
        "rfo_nthreads": (c_int64, [c_int32, c_int32, c_int32, c_int32]),
        "rfo_trace": (None, [P, c_int32, P, c_int32, POINTER(Opts), c_int32, c_int64, c_int64, P]),
        "rfo_solve": (c_int32, [P, c_int32, P, c_int32, POINTER(Opts), c_int32, c_int32, P, P, P]),
        "rfo_solve_range": (c_int32, [P, c_int32, P, c_int32, POINTER(Opts), c_int32, c_int64, c_int64, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


_L = None


def lib():
    global _L
    if _L is None:
        _L = _load()
    return _L


def _p(a):
    return None if a is None else a.ctypes.data


def philox(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().rfo_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def u01(u: int) -> float:
    return lib().rfo_u01_24(u)


def sincosf(x: float) -> tuple[float, float]:
    s, c = c_float(), c_float()
    lib().rfo_sincosf(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def sincos64(x: float) -> tuple[float, float]:
    s, c = c_double(), c_double()
    lib().rfo_sincos64(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def workitem_u32(rng: str, seed: int, geometry: int, emitter: int, wi: int) -> np.ndarray:
    o = np.zeros(4, dtype=np.uint32)
    lib().rfo_workitem_u32(RNG[rng], seed, geometry, emitter, wi, _p(o))
    return o


def emit_f32(prim, rands) -> np.ndarray:
    p = np.ascontiguousarray(np.asarray(prim, dtype=PRIM_DTYPE).reshape(1))
    r = np.ascontiguousarray(rands, dtype=np.float32)
    o = np.zeros(6, dtype=np.float32)
    lib().rfo_emit_f32(_p(p), _p(r), _p(o))
    return o


def emit_f64(prim, rands) -> np.ndarray:
    p = np.ascontiguousarray(np.asarray(prim, dtype=PRIM_DTYPE).reshape(1))
    r = np.ascontiguousarray(rands, dtype=np.float64)
    o = np.zeros(6, dtype=np.float64)
    lib().rfo_emit_f64(_p(p), _p(r), _p(o))
    return o


def hit_f32(ptype: int, invm, util1: float, util2: float, ray) -> float:
    m = np.ascontiguousarray(invm, dtype=np.float32)
    r = np.ascontiguousarray(ray, dtype=np.float32)
    return lib().rfo_hit_f32(ptype, _p(m), util1, util2, _p(r))


def hit_f64(ptype: int, invm, util1: float, util2: float, ray) -> float:
    m = np.ascontiguousarray(invm, dtype=np.float32)
    r = np.ascontiguousarray(ray, dtype=np.float64)
    return lib().rfo_hit_f64(ptype, _p(m), util1, util2, _p(r))


def find_f32(prims, start: int, end: int, ray) -> int:
    p = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
    r = np.ascontiguousarray(ray, dtype=np.float32)
    return lib().rfo_find_f32(_p(p), start, end, _p(r))


def find_f64(prims, start: int, end: int, ray) -> int:
    p = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
    r = np.ascontiguousarray(ray, dtype=np.float64)
    return lib().rfo_find_f64(_p(p), start, end, _p(r))


def find_bv_f32(prims, bvs, ray) -> int:
    p = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
    b = np.ascontiguousarray(bvs, dtype=BV_DTYPE)
    r = np.ascontiguousarray(ray, dtype=np.float32)
    return lib().rfo_find_bv_f32(_p(p), _p(b), len(b), _p(r))


def find_bv_f64(prims, bvs, ray) -> int:
    p = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
    b = np.ascontiguousarray(bvs, dtype=BV_DTYPE)
    r = np.ascontiguousarray(ray, dtype=np.float64)
    return lib().rfo_find_bv_f64(_p(p), _p(b), len(b), _p(r))


def nthreads(ray_density: int, n_samples: int, count: int = 4, wgsize: int = 128) -> int:
    return lib().rfo_nthreads(ray_density, n_samples, count, wgsize)


def _opts(mode="compat", use_bvs=False, rng="reference", seed=0, geometry=0, threads=None) -> Opts:
    o = Opts()
    o.mode = MODE[mode]
    o.use_bvs = int(bool(use_bvs))
    o.rng = RNG[rng]
    o.seed = seed & 0xFFFFFFFF
    o.geometry = geometry & 0xFFFFFFFF
    o.n_threads = threads if threads else min(16, os.cpu_count() or 1)
    return o


def solve(prims, bvs=None, n_wi=1, mode="compat", use_bvs=False, rng="reference", seed=0, geometry=0,
          threads=None, emitters=None):
    """Hit matrix (n x n int64) and rays per emitter.  n_wi: work items per
    emitter (scalar or per-emitter array); emitters: optional (e0, e1) range."""
    p = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
    n = len(p)
    b = None if bvs is None or len(bvs) == 0 else np.ascontiguousarray(bvs, dtype=BV_DTYPE)
    nw = np.broadcast_to(np.asarray(n_wi, dtype=np.int64), (n,)).copy()
    hits = np.zeros((n, n), dtype=np.int64)
    rays = np.zeros(n, dtype=np.int64)
    e0, e1 = emitters if emitters else (0, n)
    o = _opts(mode, use_bvs, rng, seed, geometry, threads)
    rc = lib().rfo_solve(_p(p), n, _p(b), 0 if b is None else len(b), ctypes.byref(o), e0, e1, _p(nw),
                         _p(hits), _p(rays))
    if rc != 0:
        raise ValueError(f"oracle rejected the scene (code {rc})")
    return hits, rays


def trace(prims, bvs, emitter: int, wi0: int, n_wi: int, mode="compat", use_bvs=False, rng="reference", seed=0,
          geometry=0) -> np.ndarray:
    p = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
    b = None if bvs is None or len(bvs) == 0 else np.ascontiguousarray(bvs, dtype=BV_DTYPE)
    out = np.zeros(4 * n_wi, dtype=np.int32)
    o = _opts(mode, use_bvs, rng, seed, geometry, 1)
    lib().rfo_trace(_p(p), len(p), _p(b), 0 if b is None else len(b), ctypes.byref(o), emitter, wi0, n_wi, _p(out))
    return out


def solve_range(prims, bvs, emitter: int, wi0: int, n_wi: int, mode="compat", use_bvs=False, rng="reference",
                seed=0, geometry=0) -> np.ndarray:
    """Hit row of one emitter over work items [wi0, wi0+n_wi) (single thread)."""
    p = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
    b = None if bvs is None or len(bvs) == 0 else np.ascontiguousarray(bvs, dtype=BV_DTYPE)
    row = np.zeros(len(p), dtype=np.int64)
    o = _opts(mode, use_bvs, rng, seed, geometry, 1)
    rc = lib().rfo_solve_range(_p(p), len(p), _p(b), 0 if b is None else len(b), ctypes.byref(o), emitter, wi0,
                               n_wi, _p(row))
    if rc != 0:
        raise ValueError(f"oracle rejected the scene (code {rc})")
    return row
