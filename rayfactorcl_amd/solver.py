"""Host-side mirror of RayFactorCL's dispatch (RayFactorCL/main.cpp:54-272).

``run_cl_gpurng`` keeps the reference's observables (int32 hit matrix,
``samplesForObj``, float view factors, ``output.txt``); ``Context`` /
``DeviceScene`` / ``solve`` expose the per-emitter launch and the whole
dispatch with int64 tallies, the batched persistent form, per-ray traces,
the autotuner and the streaming service.  Every call goes through the C ABI
of ``librayfactor_amd.so``.
"""
from __future__ import annotations

import ctypes
import weakref
from ctypes import byref, c_double, c_int32, c_int64, c_void_p
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import BV_DTYPE, FORM, MATH, PRIM_DTYPE, RNG, TRANSFORM, SolveOpts, Timing, check, lib, ptr


def make_opts(math: str = "compat", rng: str = "reference", use_bvs: bool = False,
              n_samples: int | None = None, rays_per_emitter: int | None = None,
              form: str = "persistent", refine: bool = True, seed: int = 0, geometry: int = 0,
              rays_per_thread: int = 0, block_size: int = 0, transform: str = "auto",
              refine_margin: float = 0.0, per_test: bool | int = False, count_tests: bool = True,
              band_tile: int = 0, hierarchy: bool = True) -> SolveOpts:
    """Solve options.  The ray budget is the reference's launch geometry
    (``n_samples``, main.cpp:172-173) unless ``rays_per_emitter`` is given.
    ``per_test``: False / 0 = band stacks, sorted ones searched; True / 1 =
    every cylinder test in full; 2 = band stacks with every band's z tested.
    The tallies are the same in all three.  ``count_tests``: the fast band
    search counts its evaluations (timing ``tests`` / ``flops`` are -1 without).
    ``band_tile``: LDS copy of the band records for the fast band search
    (0 = the autotuner's choice, off until tuned; -1 = off; n = at most n
    records).  ``hierarchy``: the fast form finds the per-test primitives of
    scenes with at least 48 of them through bounding-box hierarchies (False:
    every primitive tested); the same tallies either way."""
    o = SolveOpts()
    lib().rf_solve_opts_default(byref(o))
    o.math = MATH[math]
    o.rng = RNG[rng]
    o.use_bvs = int(bool(use_bvs))
    o.form = FORM[form]
    o.refine = int(bool(refine))
    o.seed = seed & 0xFFFFFFFF
    o.geometry = geometry & 0xFFFFFFFF
    o.rays_per_thread = rays_per_thread
    o.block_size = block_size
    o.transform = TRANSFORM[transform]
    o.refine_margin = float(refine_margin)
    o.per_test = int(per_test)
    o.count_tests = int(bool(count_tests))
    o.band_tile = int(band_tile)
    o.hierarchy = 0 if hierarchy else -1
    if rays_per_emitter is not None:
        o.budget = 1
        o.rays_per_emitter = int(rays_per_emitter)
    else:
        o.budget = 0
        o.n_samples = int(n_samples if n_samples is not None else 1000000)
    return o


def nthreads(ray_density: int, n_samples: int, count: int = 4, wgsize: int = 128) -> int:
    """Work items launched per emitter (main.cpp:172-173)."""
    out = c_int64()
    check(lib().rf_nthreads(ray_density, n_samples, count, wgsize, byref(out)), "rf_nthreads")
    return out.value


class Context:
    """One HIP device and stream (replaces opencl_init, util_opencl.h:133)."""

    def __init__(self, device: int = 0):
        h = c_void_p()
        check(lib().rf_ctx_create(device, byref(h)), "rf_ctx_create")
        self.handle = h
        name = ctypes.create_string_buffer(256)
        n_cu, clk = c_int32(), c_int32()
        lib().rf_device_info(self.handle, name, 256, byref(n_cu), byref(clk))
        self.arch = name.value.decode()
        self.n_cu = n_cu.value
        self.clock_mhz = clk.value
        self._children = weakref.WeakSet()   # scenes and streams, closed before the context

    def close(self) -> None:
        if self.handle:
            for child in list(self._children):
                child.close()
            lib().rf_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self) -> None:
        check(lib().rf_sync(self.handle), "rf_sync", self.handle)


class DeviceScene:
    """Primitives (and bounding volumes) resident on the device."""

    def __init__(self, ctx: Context, prims: np.ndarray, bvs: np.ndarray | None = None):
        self.ctx = ctx
        self.prims = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
        self.bvs = None if bvs is None or len(bvs) == 0 else np.ascontiguousarray(bvs, dtype=BV_DTYPE)
        h = c_void_p()
        check(lib().rf_scene_create(ctx.handle, ptr(self.prims), len(self.prims), ptr(self.bvs),
                                    0 if self.bvs is None else len(self.bvs), byref(h)),
              "rf_scene_create", ctx.handle)
        self.handle = h
        self.n = len(self.prims)
        ctx._children.add(self)

    def close(self) -> None:
        # the context releases its scenes when it closes first
        if self.handle and self.ctx.handle:
            lib().rf_scene_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- per-emitter launch (clSetKernelArg(0) + clEnqueueNDRangeKernel) --
    def launch_emitter(self, emitter: int, wi0: int, n_wi: int, opts: SolveOpts) -> None:
        check(lib().rf_launch_emitter(self.ctx.handle, self.handle, emitter, wi0, n_wi, byref(opts)),
              "rf_launch_emitter", self.ctx.handle)

    def hits(self) -> np.ndarray:
        out = np.zeros((self.n, self.n), dtype=np.int64)
        check(lib().rf_hits_read(self.ctx.handle, self.handle, ptr(out)), "rf_hits_read", self.ctx.handle)
        return out

    def vf(self, rays_per_emitter: int) -> np.ndarray:
        """View factors of the device tally, normalised on the device in FP64."""
        out = np.empty((self.n, self.n), dtype=np.float64)
        check(lib().rf_vf_read(self.ctx.handle, self.handle, int(rays_per_emitter), ptr(out)), "rf_vf_read",
              self.ctx.handle)
        return out

    def reset_hits(self) -> None:
        check(lib().rf_hits_reset(self.ctx.handle, self.handle), "rf_hits_reset", self.ctx.handle)


@dataclass
class SolveResult:
    hits: np.ndarray | None  # (n, n) int64, row = emitter (None: batch_solve without readback)
    rays: np.ndarray        # (n,) int64, rays traced per emitter
    timing: dict

    @property
    def vf(self) -> np.ndarray:
        return normalise(self.hits, self.rays)


def solve(ctx: Context, scene: DeviceScene, opts: SolveOpts | None = None, **kw) -> SolveResult:
    """All emitters of one scene (the dispatch of runCL_GPURNG, main.cpp:54-214)."""
    o = opts if opts is not None else make_opts(**kw)
    hits = np.zeros((scene.n, scene.n), dtype=np.int64)
    rays = np.zeros(scene.n, dtype=np.int64)
    t = Timing()
    check(lib().rf_solve(ctx.handle, scene.handle, byref(o), ptr(hits), ptr(rays), byref(t)), "rf_solve",
          ctx.handle)
    return SolveResult(hits, rays, t.as_dict())


def batch_solve(ctx: Context, scenes: list[DeviceScene], opts: SolveOpts | None = None, readback: bool = True,
                **kw) -> tuple[list[SolveResult], dict]:
    """Independent scenes in one persistent launch (form='persistent') or one
    launch per emitter of each (form='per_emitter' / 'reference')."""
    o = opts if opts is not None else make_opts(**kw)
    n_g = len(scenes)
    sc = (c_void_p * n_g)(*[s.handle.value for s in scenes])
    # without readback no host tally is allocated (hits of the results are None):
    # zeroing n x n int64 arrays for a whole sweep costs tens of ms per call
    hits = [np.zeros((s.n, s.n), dtype=np.int64) if readback else None for s in scenes]
    rays = [np.zeros(s.n, dtype=np.int64) for s in scenes]
    hp = (c_void_p * n_g)(*[h.ctypes.data if readback else None for h in hits])
    rp = (c_void_p * n_g)(*[r.ctypes.data for r in rays])
    t = Timing()
    check(lib().rf_batch_solve(ctx.handle, sc, n_g, byref(o), hp, rp, byref(t)), "rf_batch_solve", ctx.handle)
    res = [SolveResult(h, r, {}) for h, r in zip(hits, rays)]
    return res, t.as_dict()


def trace(ctx: Context, scene: DeviceScene, emitter: int, wi0: int, n_wi: int, opts: SolveOpts | None = None,
          **kw) -> np.ndarray:
    """Hit index (-1 = miss) of each of the 4*n_wi rays of work items [wi0, wi0+n_wi)."""
    o = opts if opts is not None else make_opts(**kw)
    out = np.empty(4 * n_wi, dtype=np.int32)
    check(lib().rf_trace_rays(ctx.handle, scene.handle, emitter, wi0, n_wi, byref(o), ptr(out)), "rf_trace_rays",
          ctx.handle)
    return out


def autotune(ctx: Context, scenes: list[DeviceScene], opts: SolveOpts | None = None, **kw) -> tuple[int, int]:
    """Fastest (rays per thread, block size) for the fast kernel on these scenes;
    rays per thread 0 means the MFMA-transform variant won."""
    o = opts if opts is not None else make_opts(**kw)
    sc = (c_void_p * len(scenes))(*[s.handle.value for s in scenes])
    r, b = c_int32(), c_int32()
    check(lib().rf_autotune(ctx.handle, sc, len(scenes), byref(o), byref(r), byref(b)), "rf_autotune", ctx.handle)
    return r.value, b.value


def run_cl_gpurng(ctx: Context, prims: np.ndarray, bvs: np.ndarray | None, use_bvs: bool, n_samples: int):
    """runCL_GPURNG (main.cpp:54-272): int32 hits, samplesForObj, float vfm."""
    prims = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
    bvs = None if bvs is None or len(bvs) == 0 else np.ascontiguousarray(bvs, dtype=BV_DTYPE)
    n = len(prims)
    hits = np.zeros((n, n), dtype=np.int32)
    samples = np.zeros(n, dtype=np.float64)
    vfm = np.zeros((n, n), dtype=np.float32)
    check(lib().rf_run_cl_gpurng(ctx.handle, ptr(prims), n, ptr(bvs), 0 if bvs is None else len(bvs),
                                 int(bool(use_bvs)), int(n_samples), ptr(hits), ptr(samples), ptr(vfm)),
          "rf_run_cl_gpurng", ctx.handle)
    return hits, samples, vfm


def normalise(hits: np.ndarray, rays: np.ndarray) -> np.ndarray:
    """FP64 view factors hits[i, j] / rays[i]."""
    hits = np.ascontiguousarray(hits, dtype=np.int64)
    rays = np.ascontiguousarray(rays, dtype=np.int64)
    n = len(rays)
    vf = np.zeros((n, n), dtype=np.float64)
    check(lib().rf_normalise(n, ptr(hits), ptr(rays), ptr(vf)), "rf_normalise")
    return vf


def selftest_math(ctx: Context) -> dict:
    """Exhaustive device check of the compat reciprocal and square root
    against IEEE 1/x and sqrt (all 2^32 inputs)."""
    mism = np.zeros(4, dtype=np.uint64)
    first = np.zeros(4, dtype=np.uint32)
    check(lib().rf_selftest_math(ctx.handle, ptr(mism), ptr(first)), "rf_selftest_math", ctx.handle)
    names = ["recip_c", "sqrt_c", "recip_newton_unguarded", "v_sqrt_unadjusted"]
    return {k: {"mismatches": int(m), "first": int(f)} for k, m, f in zip(names, mism, first)}


def selftest_sincos(ctx: Context) -> tuple[float, float]:
    """Largest absolute error of the fast emission's sin and cos of 2 pi u
    over every u = k 2^-24 (all 2^24 emission uniforms), against FP64."""
    err = np.zeros(2, dtype=np.float64)
    check(lib().rf_selftest_sincos(ctx.handle, ptr(err)), "rf_selftest_sincos", ctx.handle)
    return float(err[0]), float(err[1])


def areas(prims: np.ndarray) -> np.ndarray:
    """Surface area of each primitive (object shape under its m transform, by quadrature)."""
    prims = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
    out = np.zeros(len(prims), dtype=np.float64)
    check(lib().rf_areas(ptr(prims), len(prims), ptr(out)), "rf_areas")
    return out


def vf_report(vf: np.ndarray, area: np.ndarray, min_product: float = 0.0) -> dict:
    """Closure (row sums) and reciprocity (A_i F_ij vs A_j F_ji) of a view-factor matrix.

    Returns the report fields plus ``closure``, the per-emitter row sums.
    """
    vf = np.ascontiguousarray(vf, dtype=np.float64)
    area = np.ascontiguousarray(area, dtype=np.float64)
    n = len(area)
    if vf.shape != (n, n):
        raise ValueError(f"vf has shape {vf.shape}, expected {(n, n)}")
    closure = np.zeros(n, dtype=np.float64)
    rep = L.VfReport()
    check(lib().rf_vf_check(n, ptr(vf), ptr(area), float(min_product), ptr(closure), byref(rep)), "rf_vf_check")
    out = rep.as_dict()
    out["closure"] = closure
    return out


def write_output(path: str, hits: np.ndarray, samples: np.ndarray) -> None:
    """output.txt in the reference format (main.cpp:236-251)."""
    hits = np.ascontiguousarray(hits, dtype=np.int64)
    samples = np.ascontiguousarray(samples, dtype=np.float64)
    check(lib().rf_write_output(str(path).encode(), len(samples), ptr(hits), ptr(samples)), "rf_write_output")


class Stream:
    """Streaming solves: resident scenes, HBM-footprint admission, batched flushes."""

    def __init__(self, ctx: Context, hbm_limit_bytes: int):
        h = c_void_p()
        check(lib().rf_stream_create(ctx.handle, int(hbm_limit_bytes), byref(h)), "rf_stream_create", ctx.handle)
        self.ctx = ctx
        self.handle = h
        self._n = {}
        ctx._children.add(self)

    def submit(self, prims: np.ndarray, opts: SolveOpts) -> int:
        prims = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
        rid = c_int64()
        check(lib().rf_stream_submit(self.handle, ptr(prims), len(prims), byref(opts), byref(rid)),
              "rf_stream_submit", self.ctx.handle)
        self._n[rid.value] = len(prims)
        return rid.value

    def flush(self) -> int:
        n = c_int32()
        check(lib().rf_stream_flush(self.handle, byref(n)), "rf_stream_flush", self.ctx.handle)
        return n.value

    def step(self) -> int:
        """One launch of queued work; returns the requests it completed."""
        n = c_int32()
        check(lib().rf_stream_step(self.handle, byref(n)), "rf_stream_step", self.ctx.handle)
        return n.value

    def collect(self, rid: int) -> tuple[np.ndarray, np.ndarray, float]:
        n = self._n[rid]
        hits = np.zeros((n, n), dtype=np.int64)
        rays = np.zeros(n, dtype=np.int64)
        lat = c_double()
        check(lib().rf_stream_collect(self.handle, rid, ptr(hits), ptr(rays), byref(lat)), "rf_stream_collect",
              self.ctx.handle)
        del self._n[rid]
        return hits, rays, lat.value

    def resident_bytes(self) -> int:
        return lib().rf_stream_resident_bytes(self.handle)

    def set_cost_limits(self, max_queued_tests: float = 0.0, max_launch_tests: float = 0.0) -> None:
        """Admission limit on the queued cost and packing limit per launch, in
        estimated ray-primitive evaluations (0 = none)."""
        check(lib().rf_stream_set_cost_limits(self.handle, float(max_queued_tests), float(max_launch_tests)),
              "rf_stream_set_cost_limits", self.ctx.handle)

    def queued_cost(self) -> float:
        return lib().rf_stream_queued_cost(self.handle)

    def close(self) -> None:
        if self.handle and self.ctx.handle:
            lib().rf_stream_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def same_tallies(a: SolveOpts, b: SolveOpts) -> bool:
    """Whether two solves may share one persistent launch (every option that can
    change a tally is equal); host only."""
    r = lib().rf_opts_same_tallies(byref(a), byref(b))
    if r < 0:
        raise L.RayFactorError(r, "rf_opts_same_tallies")
    return bool(r)


def fail_solve(ctx: Context, nth: int) -> None:
    """Test hook: the nth solve issued on ctx from now fails (0 cancels)."""
    check(lib().rf_debug_fail_solve(ctx.handle, int(nth)), "rf_debug_fail_solve", ctx.handle)


def scene_footprint(n_prims: int, n_bvs: int = 0) -> int:
    return lib().rf_scene_footprint(n_prims, n_bvs)
