"""The fast form's bounding-box hierarchy over per-test primitives
(rf_solve_opts.hierarchy) against the same scan with every primitive tested:
identical tallies and identical ray-for-ray hits (the hierarchy form sorts
each tile of rays by direction before walking them), in every launch form and
shape, with FP64 re-solves on and off, with band stacks in the same scene,
with the re-solve list overflowing (re-scans), far from the origin, and
batched with scenes that have no hierarchy.  The scan with every primitive
tested is itself held to the FP64 form (test_gpu_fast) and, through it, to
the oracle."""
import numpy as np
import pytest

import rayfactorcl_amd as rf
from rayfactorcl_amd import _lib, scenes

pytestmark = pytest.mark.gpu

KW = dict(math="fast", rng="seeded", form="persistent")


def _soup(n, seed=3, shift=0.0):
    sc = scenes.parse_xml(scenes.soup_xml(n, seed=seed))
    p = sc.prims.copy()
    if shift:
        # world translation by +shift in x, y and z: invm's offsets absorb
        # -A shift, m's translation +shift
        for r in range(3):
            p["invm"][:, 4 * r + 3] -= (p["invm"][:, 4 * r:4 * r + 3] * np.float32(shift)).sum(axis=1)
            p["m"][:, 4 * r + 3] += np.float32(shift)
    return p


def _pair(ctx, prims, **kw):
    dev = rf.DeviceScene(ctx, prims, None)
    a = rf.solve(ctx, dev, hierarchy=True, **KW, **kw)
    b = rf.solve(ctx, dev, hierarchy=False, **KW, **kw)
    return dev, a, b


@pytest.mark.parametrize("refine", [True, False])
@pytest.mark.parametrize("rpt", [0, 1, 2, 4, 8])
def test_soup_tallies_equal_scan(ctx, refine, rpt):
    dev, a, b = _pair(ctx, _soup(400), rays_per_emitter=1 << 13, refine=refine, rays_per_thread=rpt)
    assert a.timing["shape_flags"] & _lib.RF_SHAPE_HIERARCHY
    assert not b.timing["shape_flags"] & _lib.RF_SHAPE_HIERARCHY
    # the default for a scene without band stacks is 4 rays per thread; 8 may
    # fall back to 4 where the traversal stacks leave no LDS
    assert a.timing["shape_rays_per_thread"] in ((rpt or 4),) + ((4,) if rpt == 8 else ())
    assert np.array_equal(a.hits, b.hits)
    # far fewer tests than the scan of every primitive
    assert 0 < a.timing["tests"] < 0.5 * b.timing["tests"]
    print(f"[hierarchy] refine={refine} rpt={rpt}: tests {a.timing['tests'] / a.timing['rays']:.1f} per ray "
          f"(scan {b.timing['tests'] / b.timing['rays']:.0f}), refined {a.timing['refined']} vs "
          f"{b.timing['refined']}", flush=True)


def test_soup_traces_equal_scan(ctx):
    prims = _soup(1000, seed=11)
    dev = rf.DeviceScene(ctx, prims, None)
    diff = hit = 0
    for e in range(0, len(prims), 97):
        a = rf.trace(ctx, dev, e, 0, 4096, math="fast", hierarchy=True)
        b = rf.trace(ctx, dev, e, 0, 4096, math="fast", hierarchy=False)
        diff += int((a != b).sum())
        hit += int((b >= 0).sum())
    assert hit > 0
    assert diff == 0


@pytest.mark.parametrize("form", ["per_emitter", "reference"])
def test_forms_equal(ctx, form):
    prims = _soup(300, seed=5)
    dev = rf.DeviceScene(ctx, prims, None)
    a = rf.solve(ctx, dev, math="fast", rng="seeded", form=form, rays_per_emitter=4096, hierarchy=True)
    b = rf.solve(ctx, dev, hierarchy=False, rays_per_emitter=4096, **KW)
    assert np.array_equal(a.hits, b.hits)


def test_with_band_stacks(ctx):
    """A furnace (band stacks, searched) with a soup inside: stacks and
    hierarchies in one scan."""
    f = scenes.parse_xml(scenes.furnace_xml(wall_bands=40, preform_bands=12)).prims
    s = _soup(200, seed=7)
    # soup centres U[0,10]^3 scaled into the furnace's interior
    prims = np.concatenate([f, s])
    dev, a, b = _pair(ctx, prims, rays_per_emitter=1 << 12)
    assert a.timing["shape_flags"] & _lib.RF_SHAPE_HIERARCHY
    assert a.timing["shape_flags"] & _lib.RF_SHAPE_STACKS
    assert np.array_equal(a.hits, b.hits)
    c = rf.solve(ctx, dev, per_test=2, rays_per_emitter=1 << 12, **KW)
    assert c.timing["shape_flags"] & _lib.RF_SHAPE_HIERARCHY
    assert np.array_equal(c.hits, b.hits)
    d = rf.solve(ctx, dev, per_test=1, rays_per_emitter=1 << 12, **KW)
    assert not d.timing["shape_flags"] & _lib.RF_SHAPE_HIERARCHY
    assert np.array_equal(d.hits, b.hits)


def test_resolve_overflow_rescans(ctx):
    """A wide re-solve margin flags many pairs per ray: the list overflows and
    lanes re-scan; the tallies stay those of the scan."""
    dev, a, b = _pair(ctx, _soup(400, seed=13), rays_per_emitter=1 << 12, refine_margin=0.25)
    assert a.timing["rescans"] > 0
    assert np.array_equal(a.hits, b.hits)


def test_far_from_origin(ctx):
    dev, a, b = _pair(ctx, _soup(300, seed=17, shift=2.0e4), rays_per_emitter=1 << 12)
    assert a.timing["shape_flags"] & _lib.RF_SHAPE_HIERARCHY
    assert np.array_equal(a.hits, b.hits)
    assert a.hits.sum() > 0


def test_batched_mixed(ctx):
    """Soups of several sizes batched: hierarchies; with a furnace (no
    hierarchy) in the batch: its units scan in the same launch.  Same
    tallies each way."""
    soups = [rf.DeviceScene(ctx, _soup(n, seed=20 + n), None) for n in (60, 150, 700)]
    fur = rf.DeviceScene(ctx, scenes.parse_xml(scenes.furnace_xml()).prims, None)
    ra, ta = rf.batch_solve(ctx, soups, hierarchy=True, rays_per_emitter=1 << 12, **KW)
    rb, tb = rf.batch_solve(ctx, soups, hierarchy=False, rays_per_emitter=1 << 12, **KW)
    assert ta["shape_flags"] & _lib.RF_SHAPE_HIERARCHY
    for x, y in zip(ra, rb):
        assert np.array_equal(x.hits, y.hits)
    # each scene's random stream follows its index in the batch
    rc, tc = rf.batch_solve(ctx, soups + [fur], rays_per_emitter=1 << 12, **KW)
    rd, td = rf.batch_solve(ctx, soups + [fur], hierarchy=False, rays_per_emitter=1 << 12, **KW)
    assert tc["shape_flags"] & _lib.RF_SHAPE_HIERARCHY
    assert not td["shape_flags"] & _lib.RF_SHAPE_HIERARCHY
    for x, y in zip(rc, rd):
        assert np.array_equal(x.hits, y.hits)
    for x, y in zip(rc, rb):
        assert np.array_equal(x.hits, y.hits)
    # the furnace's scan tests counted on the host, the soups' on the device
    assert 0 < tc["tests"] < td["tests"]


def test_uncounted(ctx):
    dev = rf.DeviceScene(ctx, _soup(200, seed=23), None)
    a = rf.solve(ctx, dev, count_tests=False, rays_per_emitter=1 << 12, **KW)
    b = rf.solve(ctx, dev, count_tests=True, rays_per_emitter=1 << 12, **KW)
    assert a.timing["tests"] == -1 and a.timing["flops"] == -1
    assert b.timing["tests"] > 0 and b.timing["flops"] > 0
    assert np.array_equal(a.hits, b.hits)


def test_small_scene_has_none(ctx):
    """Fewer than 48 primitives outside band stacks: no hierarchy."""
    dev = rf.DeviceScene(ctx, _soup(40, seed=29), None)
    a = rf.solve(ctx, dev, rays_per_emitter=1 << 10, **KW)
    assert not a.timing["shape_flags"] & _lib.RF_SHAPE_HIERARCHY


def test_large_scene_keeps_hierarchy(ctx):
    """20 000 primitives: the hierarchy's depth sizes the lanes' LDS stacks;
    traces equal the scan of every primitive."""
    prims = _soup(20000, seed=37)
    dev = rf.DeviceScene(ctx, prims, None)
    _, t = rf.batch_solve(ctx, [dev], readback=False, rays_per_emitter=8, **KW)
    assert t["shape_flags"] & _lib.RF_SHAPE_HIERARCHY
    diff = hit = 0
    for e in (0, 4999, 13000, 19999):
        a = rf.trace(ctx, dev, e, 0, 2048, math="fast", hierarchy=True)
        b = rf.trace(ctx, dev, e, 0, 2048, math="fast", hierarchy=False)
        diff += int((a != b).sum())
        hit += int((b >= 0).sum())
    assert hit > 0 and diff == 0


def test_config3_scene_equal_scan(ctx):
    """The 1000-primitive soup of configuration 3 at 2^16 rays per emitter
    (6.6e7 rays): identical tallies with and without the hierarchy, FP64
    re-solves on."""
    prims = scenes.parse_xml(scenes.soup_xml(1000, seed=0x52465331)).prims
    dev, a, b = _pair(ctx, prims, rays_per_emitter=1 << 16)
    assert a.timing["shape_flags"] & _lib.RF_SHAPE_HIERARCHY
    assert a.timing["refined"] > 0
    assert np.array_equal(a.hits, b.hits)
    print(f"[hierarchy] config-3 soup: {a.timing['rays'] / a.timing['kernel_ms'] * 1e3:.3e} rays/s against "
          f"{b.timing['rays'] / b.timing['kernel_ms'] * 1e3:.3e} for the full scan", flush=True)


def _box_zoo(n_rect=60, n_soup=150, seed=43):
    """A closed box [0,10]^3 of six axis-aligned rectangles (edges shared, so
    rays meet exact ties where faces meet), random rectangles inside, and a
    soup: every leaf type, axis-parallel rays' slabs, and ties."""
    P = scenes._prim
    rng = np.random.default_rng(seed)
    body = [
        P("rectangle", scale=(5.0, 5.0, 1.0), translate=(5.0, 5.0, 0.0)),
        P("rectangle", scale=(5.0, 5.0, 1.0), rotate=scenes.FLIP_X, translate=(5.0, 5.0, 10.0)),
        P("rectangle", scale=(5.0, 5.0, 1.0), rotate=(90.0, (0.0, 1.0, 0.0)), translate=(0.0, 5.0, 5.0)),
        P("rectangle", scale=(5.0, 5.0, 1.0), rotate=(-90.0, (0.0, 1.0, 0.0)), translate=(10.0, 5.0, 5.0)),
        P("rectangle", scale=(5.0, 5.0, 1.0), rotate=(-90.0, (1.0, 0.0, 0.0)), translate=(5.0, 0.0, 5.0)),
        P("rectangle", scale=(5.0, 5.0, 1.0), rotate=(90.0, (1.0, 0.0, 0.0)), translate=(5.0, 10.0, 5.0)),
    ]
    for _ in range(n_rect):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        body.append(P("rectangle", scale=(float(rng.uniform(0.2, 1.0)), float(rng.uniform(0.2, 1.0)), 1.0),
                      rotate=(float(rng.uniform(0.0, 360.0)), tuple(float(x) for x in v)),
                      translate=tuple(float(x) for x in rng.uniform(1.0, 9.0, 3))))
    box = scenes.parse_xml(scenes._doc("box zoo", "".join(body))).prims
    return np.concatenate([box, _soup(n_soup, seed=seed)])


@pytest.mark.parametrize("refine", [True, False])
def test_box_zoo_equal_scan(ctx, refine):
    prims = _box_zoo()
    dev, a, b = _pair(ctx, prims, rays_per_emitter=1 << 13, refine=refine)
    assert a.timing["shape_flags"] & _lib.RF_SHAPE_HIERARCHY
    assert np.array_equal(a.hits, b.hits)
    # the box closes the scene: almost every ray hits something
    assert a.hits.sum() > 0.95 * a.rays.sum()
    for e in (0, 3, 10, 100, len(prims) - 1):
        ta = rf.trace(ctx, dev, e, 0, 2048, math="fast", hierarchy=True, refine=refine)
        tb = rf.trace(ctx, dev, e, 0, 2048, math="fast", hierarchy=False, refine=refine)
        assert np.array_equal(ta, tb), e


def test_fp64_walk_equal_scan(ctx):
    """The FP64 form walks the hierarchy where it holds every primitive (no
    band stacks): the same tallies as its scan; with band stacks in the
    scene it scans."""
    dev = rf.DeviceScene(ctx, _soup(400, seed=47), None)
    kw = dict(math="fp64", rng="seeded", form="persistent", rays_per_emitter=1 << 12)
    a = rf.solve(ctx, dev, hierarchy=True, **kw)
    b = rf.solve(ctx, dev, hierarchy=False, **kw)
    assert a.timing["shape_flags"] & _lib.RF_SHAPE_HIERARCHY
    assert not b.timing["shape_flags"] & _lib.RF_SHAPE_HIERARCHY
    assert np.array_equal(a.hits, b.hits)
    assert 0 < a.timing["tests"] < 0.5 * b.timing["tests"]
    f = scenes.parse_xml(scenes.furnace_xml(wall_bands=40, preform_bands=12)).prims
    dev2 = rf.DeviceScene(ctx, np.concatenate([f, _soup(200, seed=7)]), None)
    c = rf.solve(ctx, dev2, hierarchy=True, **kw)
    d = rf.solve(ctx, dev2, hierarchy=False, **kw)
    assert not c.timing["shape_flags"] & _lib.RF_SHAPE_HIERARCHY
    assert np.array_equal(c.hits, d.hits)


def test_fp64_walk_traces_match_oracle(ctx, oracle):
    """FP64 traces through the hierarchy, ray for ray against the CPU oracle's
    FP64 form (a scan of every primitive in the caller's order), on the box
    zoo: ties where the box's faces meet go to the lower index in both."""
    prims = _box_zoo(n_rect=40, n_soup=100, seed=53)
    dev = rf.DeviceScene(ctx, prims, None)
    for e in (0, 5, 60, len(prims) - 1):
        a = rf.trace(ctx, dev, e, 0, 1024, math="fp64", rng="seeded")
        b = oracle.trace(prims, None, e, 0, 1024, mode="fp64", rng="seeded")
        assert np.array_equal(a, b), e
