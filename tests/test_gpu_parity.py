"""GPU parity: the HIP kernels against the CPU oracle on identical Philox
streams.  Compat and FP64 are bit-exact (integer hit matrices and per-ray
traces); every dispatch form gives the same tallies."""
import numpy as np
import pytest

import rayfactorcl_amd as rf
from conftest import SCENE_NAMES

pytestmark = pytest.mark.gpu


def wi_for(sc):
    return 4096 if sc.n <= 8 else (1024 if sc.n <= 64 else 128)


def test_compat_math_exhaustive(ctx):
    """The compat kernels reach IEEE 1/x and sqrt through short hardware
    sequences with an IEEE fallback; over all 2^32 inputs they must give the
    compiler's IEEE results bit for bit (NaN for NaN)."""
    r = rf.selftest_math(ctx)
    print(r, flush=True)
    assert r["recip_c"]["mismatches"] == 0, r
    assert r["sqrt_c"]["mismatches"] == 0, r


def test_fast_sincos_error_bound(ctx):
    """The fast emission's sin and cos of 2 pi u come from v_sin_f32 /
    v_cos_f32 (argument in revolutions).  Over all 2^24 emission uniforms
    their largest error against FP64 is about 1 ulp at 1: at most 2^-22
    (the sampled directions move by no more than that)."""
    es, ec = rf.selftest_sincos(ctx)
    print(f"[sincos] max abs error sin {es:.3e} cos {ec:.3e}", flush=True)
    assert 0 < es <= 2.0 ** -22 and 0 < ec <= 2.0 ** -22


@pytest.mark.parametrize("name", SCENE_NAMES)
@pytest.mark.parametrize("math", ["compat", "fp64"])
def test_hits_match_oracle(ctx, oracle, golden_scene, name, math):
    sc = golden_scene(name)
    n_wi = wi_for(sc)
    print(f"[parity] {name} math={math} n={sc.n} n_wi={n_wi}", flush=True)
    dev = rf.DeviceScene(ctx, sc.prims, sc.bvs)
    got = rf.solve(ctx, dev, math=math, use_bvs=sc.useBoundingVolumes, rays_per_emitter=4 * n_wi)
    want, rays = oracle.solve(sc.prims, sc.bvs, n_wi=n_wi, mode=math, use_bvs=sc.useBoundingVolumes)
    assert np.array_equal(got.rays, rays)
    diff = np.argwhere(got.hits != want)
    assert diff.size == 0, f"{len(diff)} cells differ, first {diff[:5].tolist()}"


@pytest.mark.parametrize("form", ["per_emitter", "persistent", "reference"])
@pytest.mark.parametrize("name", ["config1", "input-C14-BV", "input-W100"])
def test_forms_agree(ctx, oracle, golden_scene, name, form):
    sc = golden_scene(name)
    n_wi = wi_for(sc)
    print(f"[forms] {name} form={form}", flush=True)
    dev = rf.DeviceScene(ctx, sc.prims, sc.bvs)
    got = rf.solve(ctx, dev, math="compat", use_bvs=sc.useBoundingVolumes, rays_per_emitter=4 * n_wi, form=form)
    want, _ = oracle.solve(sc.prims, sc.bvs, n_wi=n_wi, mode="compat", use_bvs=sc.useBoundingVolumes)
    assert np.array_equal(got.hits, want)


def test_reference_budget(ctx, oracle, golden_scene):
    # launch geometry of main.cpp:172-173 at a small n_samples, every emitter with primitive 0's count
    sc = golden_scene("input-C47")
    dev = rf.DeviceScene(ctx, sc.prims, sc.bvs)
    got = rf.solve(ctx, dev, math="compat", n_samples=20000)
    nt = rf.nthreads(int(sc.prims[0]["rayDensity"]), 20000)
    assert list(got.rays) == [4 * nt] * sc.n
    want, _ = oracle.solve(sc.prims, None, n_wi=nt, mode="compat")
    assert np.array_equal(got.hits, want)


@pytest.mark.parametrize("math", ["compat", "fp64"])
def test_trace_matches_oracle(ctx, oracle, golden_scene, math):
    sc = golden_scene("config1")
    dev = rf.DeviceScene(ctx, sc.prims, sc.bvs)
    for e in range(sc.n):
        got = rf.trace(ctx, dev, e, 1000, 3000, math=math)
        want = oracle.trace(sc.prims, None, e, 1000, 3000, mode=math)
        assert np.array_equal(got, want), (e, np.argwhere(got != want)[:5])


def test_seeded_stream(ctx, oracle, golden_scene):
    sc = golden_scene("config1")
    dev = rf.DeviceScene(ctx, sc.prims, sc.bvs)
    got = rf.solve(ctx, dev, math="compat", rng="seeded", seed=1234, geometry=7, rays_per_emitter=8192)
    want, _ = oracle.solve(sc.prims, None, n_wi=2048, mode="compat", rng="seeded", seed=1234, geometry=7)
    assert np.array_equal(got.hits, want)


def test_run_cl_gpurng_observables(ctx, oracle, golden_scene, tmp_path):
    sc = golden_scene("input-C14")
    hits, samples, vfm = rf.run_cl_gpurng(ctx, sc.prims, None, False, 40000)
    nt = rf.nthreads(int(sc.prims[0]["rayDensity"]), 40000)
    assert hits.dtype == np.int32 and list(samples) == [4.0 * nt] * 2
    want, _ = oracle.solve(sc.prims, None, n_wi=nt, mode="compat")
    assert np.array_equal(hits, want)
    assert np.array_equal(vfm, (hits.astype(np.float32) / samples[:, None]).astype(np.float32))
