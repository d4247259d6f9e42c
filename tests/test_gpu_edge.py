"""Edge cases and error paths of the GPU dispatch."""
import numpy as np
import pytest

import rayfactorcl_amd as rf
from rayfactorcl_amd import _lib, scenes

pytestmark = pytest.mark.gpu


def test_single_primitive_scene(ctx, oracle, golden_scene):
    sc = golden_scene("input-uni")
    dev = rf.DeviceScene(ctx, sc.prims, None)
    r = rf.solve(ctx, dev, math="compat", rays_per_emitter=4096)
    want, _ = oracle.solve(sc.prims, None, n_wi=1024, mode="compat")
    assert np.array_equal(r.hits, want)


def test_zero_work_items(ctx, golden_scene):
    sc = golden_scene("input-C14")
    dev = rf.DeviceScene(ctx, sc.prims, None)
    r = rf.solve(ctx, dev, math="compat", n_samples=64)     # 4*64/4 = 64 -> rounds down to 0 (Q5)
    assert list(r.rays) == [0, 0] and r.hits.sum() == 0


def test_invalid_scenes_are_rejected(ctx, golden_scene):
    sc = golden_scene("input-C14")
    bad = sc.prims.copy()
    bad[1]["type"] = rf.NONE
    with pytest.raises(rf.RayFactorError) as e:
        rf.DeviceScene(ctx, bad, None)
    assert e.value.code == _lib.RF_ERR_INVALID_SCENE
    bvs = golden_scene("input-C14-BV").bvs.copy()
    bvs[1]["objEndIndex"] = 5
    with pytest.raises(rf.RayFactorError) as e:
        rf.DeviceScene(ctx, sc.prims, bvs)
    assert e.value.code == _lib.RF_ERR_INVALID_SCENE
    with pytest.raises(rf.RayFactorError):
        rf.DeviceScene(ctx, sc.prims[:0], None)


def test_bad_options_are_rejected(ctx, golden_scene):
    sc = golden_scene("input-C14")
    dev = rf.DeviceScene(ctx, sc.prims, None)
    with pytest.raises(rf.RayFactorError) as e:
        rf.solve(ctx, dev, math="fast", rays_per_emitter=4096, rays_per_thread=3)
    assert e.value.code == _lib.RF_ERR_INVALID_ARG
    with pytest.raises(rf.RayFactorError) as e:
        rf.solve(ctx, dev, math="compat", rays_per_emitter=4 * (1 << 32) + 8)   # reference RNG: 32-bit gid
    assert e.value.code == _lib.RF_ERR_OVERFLOW
    with pytest.raises(rf.RayFactorError) as e:
        rf.solve(ctx, dev, math="compat", n_samples=2 ** 30)                    # rayDensity*n_samples > int32
    assert e.value.code == _lib.RF_ERR_OVERFLOW
    for bad in (dict(refine_margin=0.5), dict(refine_margin=-1e-3), dict(refine_margin=float("nan"))):
        with pytest.raises(rf.RayFactorError) as e:
            rf.solve(ctx, dev, math="fast", rays_per_emitter=4096, **bad)
        assert e.value.code == _lib.RF_ERR_INVALID_ARG
    o = rf.make_opts(math="fast", rays_per_emitter=4096)
    o.transform = 7
    with pytest.raises(rf.RayFactorError) as e:
        rf.solve(ctx, dev, o)
    assert e.value.code == _lib.RF_ERR_INVALID_ARG
    o = rf.make_opts(math="fast", rays_per_emitter=4096)
    o.per_test = 3
    with pytest.raises(rf.RayFactorError) as e:
        rf.solve(ctx, dev, o)
    assert e.value.code == _lib.RF_ERR_INVALID_ARG


@pytest.mark.parametrize("transform", ["valu", "mfma"])
def test_large_scene_global_atomic_tally(ctx, transform):
    # more primitives than LDS bins: per-ray global atomics; rows agree with traces
    sc = scenes.parse_xml(scenes.furnace_xml(wall_bands=16500, preform_bands=10))
    assert sc.n > 16000
    dev = rf.DeviceScene(ctx, sc.prims, None)
    r = rf.solve(ctx, dev, math="fast", rays_per_emitter=256, rng="seeded", form="persistent", transform=transform)
    for e in [0, 8000, sc.n - 1]:
        tr = rf.trace(ctx, dev, e, 0, 64, math="fast", rng="seeded", transform=transform)
        row = np.bincount(tr[tr >= 0], minlength=sc.n)
        assert np.array_equal(row, r.hits[e])


def test_launch_emitter_accumulates(ctx, oracle, golden_scene):
    sc = golden_scene("config1")
    dev = rf.DeviceScene(ctx, sc.prims, None)
    o = rf.make_opts(math="compat", rays_per_emitter=0)
    dev.reset_hits()
    dev.launch_emitter(2, 0, 1000, o)
    dev.launch_emitter(2, 1000, 3000, o)
    dev.launch_emitter(0, 500, 700, o)
    ctx.sync()
    h = dev.hits()
    row2 = oracle.solve_range(sc.prims, None, 2, 0, 4000, mode="compat")
    row0 = oracle.solve_range(sc.prims, None, 0, 500, 700, mode="compat")
    assert np.array_equal(h[2], row2) and np.array_equal(h[0], row0) and h[[1, 3, 4]].sum() == 0


@pytest.mark.parametrize("form", ["per_emitter", "reference"])
def test_launch_emitter_fan_out_and_implicit_join(ctx, form):
    """Per-emitter launches of the queue kernel run on several streams; reading
    the tally (without rf_sync) waits for all of them.  The reference shape
    keeps one in-order queue.  Both equal one persistent solve."""
    sc = scenes.parse_xml(scenes.furnace_xml(wall_bands=30, preform_bands=10))
    dev = rf.DeviceScene(ctx, sc.prims, None)
    n_wi = 3000
    o = rf.make_opts(math="fast", rng="seeded", rays_per_emitter=0, form=form)
    want = rf.solve(ctx, dev, math="fast", rng="seeded", rays_per_emitter=4 * n_wi, form="persistent").hits
    for rep in range(2):
        dev.reset_hits()
        for e in range(sc.n):
            dev.launch_emitter(e, 0, n_wi // 2, o)
        for e in reversed(range(sc.n)):
            dev.launch_emitter(e, n_wi // 2, n_wi - n_wi // 2, o)
        assert np.array_equal(dev.hits(), want), rep


def test_device_fp64_normalisation_matches_host(ctx, golden_scene):
    """rf_vf_read divides the device tally by the emitters' ray count in FP64 on
    the GPU: bit-identical to the host FP64 normalisation of the same counts,
    and the rows of the closed enclosure sum to 1 less the seam losses."""
    sc = golden_scene("config1")
    dev = rf.DeviceScene(ctx, sc.prims, None)
    rays = 1 << 20
    r = rf.solve(ctx, dev, math="fast", rng="seeded", rays_per_emitter=rays)
    vf = dev.vf(rays)
    assert vf.dtype == np.float64
    assert np.array_equal(vf, rf.normalise(r.hits, r.rays))
    assert np.all(np.abs(vf[:3].sum(axis=1) - 1.0) < 2e-3)
    assert np.array_equal(dev.vf(0), np.zeros_like(vf))
    dev.close()
