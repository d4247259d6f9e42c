"""Batched persistent solves: identical tallies to one-geometry solves on the
same streams, for ragged geometry sizes, in every arithmetic form."""
import numpy as np
import pytest

import rayfactorcl_amd as rf
from rayfactorcl_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sweep8(ctx):
    geoms = scenes.sweep(8, seed=17)
    return geoms, [rf.DeviceScene(ctx, g.prims, None) for g in geoms]


@pytest.mark.parametrize("math", ["fast", "fp64"])
def test_batched_equals_single(ctx, sweep8, math):
    geoms, dev = sweep8
    rays = 4000
    res, t = rf.batch_solve(ctx, dev, math=math, rng="seeded", seed=5, geometry=100, rays_per_emitter=rays,
                            form="persistent")
    assert t["launches"] == 1 and t["rays"] == sum(4 * ((rays + 3) // 4) * g.n for g in geoms)
    for g, (d, r) in enumerate(zip(dev, res)):
        one = rf.solve(ctx, d, math=math, rng="seeded", seed=5, geometry=100 + g, rays_per_emitter=rays,
                       form="per_emitter")
        assert np.array_equal(r.hits, one.hits), g
        assert np.array_equal(r.rays, one.rays)


def test_batched_compat_matches_oracle(ctx, oracle, sweep8):
    geoms, dev = sweep8
    sub = [0, 3]
    res, _ = rf.batch_solve(ctx, [dev[i] for i in sub], math="compat", rng="seeded", seed=2, geometry=0,
                            rays_per_emitter=256, form="persistent")
    for k, i in enumerate(sub):
        want, _ = oracle.solve(geoms[i].prims, None, n_wi=64, mode="compat", rng="seeded", seed=2, geometry=k)
        assert np.array_equal(res[k].hits, want)


def test_batched_per_emitter_form(ctx, sweep8):
    geoms, dev = sweep8
    a, ta = rf.batch_solve(ctx, dev[:3], math="fast", rng="seeded", rays_per_emitter=3000, form="persistent")
    b, tb = rf.batch_solve(ctx, dev[:3], math="fast", rng="seeded", rays_per_emitter=3000, form="per_emitter")
    assert tb["launches"] == sum(g.n for g in geoms[:3])
    for x, y in zip(a, b):
        assert np.array_equal(x.hits, y.hits)


def test_closure_of_enclosure_rows(ctx, sweep8):
    # wall bands of a furnace see only the enclosure: every wall ray hits something
    # (the bottom annulus leaves the preform hole open, so allow a small escape)
    geoms, dev = sweep8
    res, _ = rf.batch_solve(ctx, dev[:2], math="fast", rng="seeded", rays_per_emitter=20000, form="persistent")
    for r in res:
        frac = r.hits.sum(axis=1) / r.rays
        assert np.all(frac <= 1.0) and np.median(frac) > 0.8


def test_graph_replays_match_eager(ctx, sweep8):
    """A persistent solve repeated with the same scenes and options is captured
    in a HIP graph on the second call and replayed afterwards.  Every replay,
    with other eager work and synchronisations in between, matches the eager
    result; a different solve in between drops the graph."""
    geoms, dev = sweep8
    sub = dev[:4]
    o = rf.make_opts(math="fast", rng="seeded", seed=11, rays_per_emitter=3000, form="persistent")
    first, t0 = rf.batch_solve(ctx, sub, o)
    assert t0["graph_launches"] == 0
    for k in range(4):
        res, t = rf.batch_solve(ctx, sub, o)
        assert t["graph_launches"] == 1 and t["launches"] == 1
        for a, b in zip(first, res):
            assert np.array_equal(a.hits, b.hits), k
        # eager work between replays
        rf.trace(ctx, sub[0], 0, 0, 256, math="fast", rng="seeded")
        rf.solve(ctx, sub[1], math="compat", rays_per_emitter=512, form="per_emitter")
        ctx.sync()
    other, t2 = rf.batch_solve(ctx, sub, rf.make_opts(math="fast", rng="seeded", seed=12, rays_per_emitter=3000,
                                                      form="persistent"))
    assert t2["graph_launches"] == 0
    again, t3 = rf.batch_solve(ctx, sub, o)
    assert t3["graph_launches"] == 0
    for a, b in zip(first, again):
        assert np.array_equal(a.hits, b.hits)
