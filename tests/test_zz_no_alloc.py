"""The allocation-free solve (rf_solve_noalloc) through the C ABI: the same
tallies as rf_solve with the same options, on consecutive calls that share
the device scratch (each smaller than the one before), into caller-owned
device memory that starts with garbage; and the capacity error."""
import ctypes
from ctypes import byref, c_size_t, c_void_p

import numpy as np
import pytest

import rayfactorcl_amd as rf
from rayfactorcl_amd import _lib, scenes

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the library uses, for the caller's device buffer."""
    h = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(c_void_p), c_size_t]
    h.hipFree.argtypes = [c_void_p]
    h.hipMemset.argtypes = [c_void_p, ctypes.c_int, c_size_t]
    h.hipMemcpy.argtypes = [c_void_p, c_void_p, c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _noalloc(ctx, dev, n, opts):
    hip = _hip()
    nbytes = n * n * 8
    buf = c_void_p()
    assert hip.hipMalloc(byref(buf), nbytes) == 0
    try:
        assert hip.hipMemset(buf, 0xA5, nbytes) == 0                  # garbage: the call clears it
        assert hip.hipDeviceSynchronize() == 0
        rc = rf.lib().rf_solve_noalloc(ctx.handle, dev.handle, byref(opts), buf.value)
        _lib.check(rc, "rf_solve_noalloc", ctx.handle)
        ctx.sync()
        out = np.empty((n, n), dtype=np.int64)
        assert hip.hipMemcpy(out.ctypes.data, buf, nbytes, 2) == 0     # device to host
    finally:
        hip.hipFree(buf)
    return out


def test_noalloc_matches_solve_on_consecutive_calls(ctx, golden_scene):
    big = scenes.sweep(8, seed=0x52465331)
    big = max(big, key=lambda g: g.n)                                   # searched band stacks
    cases = [
        # random soups: the hierarchy form, fast and FP64
        (scenes.parse_xml(scenes.soup_xml(700, seed=61)), None, dict(math="fast", rng="seeded", rays_per_emitter=4000)),
        (scenes.parse_xml(scenes.soup_xml(600, seed=62)), None, dict(math="fp64", rng="seeded", rays_per_emitter=2000)),
        (big, None, dict(math="fast", rng="seeded", rays_per_emitter=20000, seed=3)),
        (scenes.parse_xml(scenes.furnace_xml(with_bv=True)), "bv",
         dict(math="fast", rng="seeded", rays_per_emitter=12000, use_bvs=True)),
        (golden_scene("input-W100"), None, dict(math="compat", n_samples=4000)),    # reference budget
        (scenes.parse_xml(scenes.config1_xml()), None, dict(math="fast", rays_per_emitter=50000, transform="mfma")),
    ]
    sizes = [sc.n for sc, _, _ in cases]
    assert sizes == sorted(sizes, reverse=True) and len(set(sizes)) == len(sizes)
    for sc, bv, kw in cases:
        dev = rf.DeviceScene(ctx, sc.prims, sc.bvs if bv else None)
        opts = rf.make_opts(form="persistent", **kw)
        print(f"[noalloc] n={sc.n} {kw}", flush=True)
        got = _noalloc(ctx, dev, sc.n, opts)
        want = rf.solve(ctx, dev, opts).hits
        assert np.array_equal(got, want), (sc.n, kw)
        assert want.sum() > 0
        dev.close()


def test_noalloc_rejects_what_it_cannot_run(ctx):
    sc = scenes.parse_xml(scenes.soup_xml(2048, seed=5))
    dev = rf.DeviceScene(ctx, sc.prims, None)
    # 2^50 rays per emitter: over 2^31 work units for the 32-bit queue counter
    o = rf.make_opts(math="fast", rng="seeded", rays_per_emitter=1 << 50)
    rc = rf.lib().rf_solve_noalloc(ctx.handle, dev.handle, byref(o), 1 << 20)   # never written: rejected first
    assert rc == _lib.RF_ERR_CAPACITY
    assert b"2^31-1 work units" in rf.lib().rf_last_error(ctx.handle)
    o = rf.make_opts(math="fast", rays_per_emitter=4096, form="per_emitter")
    assert rf.lib().rf_solve_noalloc(ctx.handle, dev.handle, byref(o), 1 << 20) == _lib.RF_ERR_INVALID_ARG
    assert rf.lib().rf_solve_noalloc(ctx.handle, dev.handle, byref(o), None) == _lib.RF_ERR_INVALID_ARG
    dev.close()
