"""Band stacks of the fast form: consecutive cylinders whose transforms differ
only in the z offset (the bands of one coaxial wall) share one quadric solve
per ray.  Every per-band decision is the per-test scan's on that band's own
record, so the tallies, the FP64 re-solve counts and the per-ray hit indices
must equal the per-test form's exactly, for every launch shape, with and
without flag-list overflow, and on stacks that overlap, leave gaps, hold a
single band or list their bands in any order.  Both stacked forms are
checked: per_test=0 searches stacks whose z offsets are sorted (walking the
others), per_test=2 tests every band's z range.  The compat and FP64 forms
stack the caller's order and must also equal the CPU oracle's per-test scan
bit for bit."""
import numpy as np
import pytest

import rayfactorcl_amd as rf
from rayfactorcl_amd import scenes

pytestmark = pytest.mark.gpu


STACKED = [0, 2]     # per_test values of the two stacked forms: search, every band


def _pair(ctx, sc, stacked=0, **kw):
    dev = rf.DeviceScene(ctx, sc.prims, None)
    a = rf.solve(ctx, dev, math="fast", rng="seeded", per_test=True, **kw)
    b = rf.solve(ctx, dev, math="fast", rng="seeded", per_test=stacked, **kw)
    dev.close()
    return a, b


def _same(a, b):
    assert np.array_equal(a.hits, b.hits)
    assert a.timing["refined"] == b.timing["refined"]


@pytest.mark.parametrize("stacked", STACKED)
@pytest.mark.parametrize("form", ["persistent", "per_emitter", "reference"])
@pytest.mark.parametrize("rpt", [1, 2, 4, 8])
def test_stacks_equal_per_test(ctx, form, rpt, stacked):
    sc = scenes.parse_xml(scenes.furnace_xml(wall_bands=40, preform_bands=12))
    a, b = _pair(ctx, sc, stacked, rays_per_emitter=20000, form=form, rays_per_thread=rpt)
    _same(a, b)
    # the same rays; the stacked form runs fewer algorithmic flops
    assert b.timing["tests_brute_force"] == a.timing["tests_brute_force"] == a.timing["tests"]
    assert b.timing["flops"] < a.timing["flops"]


@pytest.mark.parametrize("stacked", STACKED)
@pytest.mark.parametrize("rpt", [1, 4, 8])
def test_stacks_wide_margin_overflow(ctx, rpt, stacked):
    """tau = 1/4 flags most rim crossings: re-solve lists overflow inside
    stacks, the stack's queued entries are given back and the re-scan starts
    at its first band; every flagged pair is still re-solved exactly once."""
    sc = scenes.parse_xml(scenes.furnace_xml(wall_bands=40, preform_bands=12))
    a, b = _pair(ctx, sc, stacked, rays_per_emitter=20000, form="persistent", rays_per_thread=rpt,
                 refine_margin=0.25)
    _same(a, b)
    assert b.timing["rescans"] > 0


@pytest.mark.parametrize("stacked", STACKED)
def test_stacks_without_refine(ctx, stacked):
    sc = scenes.parse_xml(scenes.furnace_xml(wall_bands=60, preform_bands=20))
    a, b = _pair(ctx, sc, stacked, rays_per_emitter=1 << 16, form="persistent", refine=False)
    assert np.array_equal(a.hits, b.hits)


def test_stacks_sweep_batched(ctx):
    geoms = scenes.sweep(6, seed=0x52465331)
    dev = [rf.DeviceScene(ctx, g.prims, None) for g in geoms]
    kw = dict(math="fast", rng="seeded", rays_per_emitter=1 << 14, form="persistent")
    a, ta = rf.batch_solve(ctx, dev, rf.make_opts(per_test=True, **kw))
    for stacked in STACKED:
        b, tb = rf.batch_solve(ctx, dev, rf.make_opts(per_test=stacked, **kw))
        for x, y in zip(a, b):
            assert np.array_equal(x.hits, y.hits)
        assert ta["refined"] == tb["refined"] and ta["tests_brute_force"] == tb["tests_brute_force"]
        # the search looks at a few bands per root; the walk at every band,
        # plus one quadric per stack (more evaluations than brute force)
        assert ta["tests"] == ta["tests_brute_force"]
        assert (tb["tests"] < ta["tests"]) if stacked == 0 else (tb["tests"] > ta["tests"])
        print(f"[stacks] sweep sample: flops per launch {ta['flops']:.3e} per test, {tb['flops']:.3e} "
              f"stacked ({stacked})", flush=True)
    for d in dev:
        d.close()


def test_stacks_traces(ctx):
    sc = scenes.parse_xml(scenes.furnace_xml(wall_bands=30, preform_bands=10))
    dev = rf.DeviceScene(ctx, sc.prims, None)
    for e in (0, 15, sc.n - 1):
        a = rf.trace(ctx, dev, e, 0, 4096, math="fast", rng="seeded", per_test=True)
        for stacked in STACKED:
            b = rf.trace(ctx, dev, e, 0, 4096, math="fast", rng="seeded", per_test=stacked)
            assert np.array_equal(a, b)
    dev.close()


def _band(lo, hi, r=0.5):
    return scenes._prim("cylinderSurface", bounding=True, small=1.0, scale=(r, r, hi - lo), translate=(0, 0, lo))


def _edge_scene():
    """Cylinder runs of every kind, in the caller's order:
    - equal-height bands with a gap (a stack whose bands do not tile);
    - a band listed twice (a stack with exactly tied bands: the first wins);
    - two overlapping bands of equal height (a stack whose z ranges overlap);
    - a run of shorter bands (a stack of its own), split by a disc in the
      caller's order (the cylinder segment keeps them consecutive);
    - single bands of other radii and heights (per-test runs);
    - bands listed top down (z offsets increasing over the stack);
    - bands listed out of z order (a stack whose offsets are not sorted:
      walked band by band in every stacked form);
    closed by a disc at each end.  All stacks but the last two list their
    bands bottom up (z offsets decreasing)."""
    parts = [_band(0.0, 0.5), _band(0.5, 1.0), _band(1.1, 1.6), _band(1.6, 2.1),
             _band(2.1, 2.6), _band(2.1, 2.6),
             _band(2.6, 3.1), _band(2.85, 3.35),
             _band(3.35, 3.6),
             scenes._prim("disc", scale=(0.3, 0.3, 1), translate=(0, 0, 1.0)),
             _band(3.6, 3.85), _band(3.85, 4.1),
             _band(0.0, 0.7, r=0.2), _band(0.7, 1.2, r=0.25), _band(4.1, 4.3, r=0.5),
             _band(2.5, 2.8, r=0.35), _band(2.2, 2.5, r=0.35), _band(1.9, 2.2, r=0.35),
             _band(3.0, 3.3, r=0.15), _band(3.6, 3.9, r=0.15), _band(3.3, 3.6, r=0.15),
             scenes._prim("disc", scale=(0.5, 0.5, 1)),
             scenes._prim("disc", scale=(0.5, 0.5, 1), rotate=scenes.FLIP_X, translate=(0, 0, 4.3))]
    return scenes.parse_xml(scenes._doc("band stack edge cases", "".join(parts)))


@pytest.mark.parametrize("stacked", STACKED)
@pytest.mark.parametrize("margin", [0.0, 0.25])
@pytest.mark.parametrize("rpt", [1, 4])
def test_stacks_edge_cases(ctx, margin, rpt, stacked):
    sc = _edge_scene()
    a, b = _pair(ctx, sc, stacked, rays_per_emitter=1 << 15, form="persistent", rays_per_thread=rpt,
                 refine_margin=margin)
    _same(a, b)
    # the duplicated band never wins over its first copy
    assert a.hits[:, 5].sum() == 0 and a.hits[:, 4].sum() > 0
    # the top-down and unsorted stacks are hit
    assert a.hits[:, 15:21].sum(axis=0).min() > 0
    dev = rf.DeviceScene(ctx, sc.prims, None)
    for e in (0, 4, 7, 12, 16, 19):
        kw = dict(math="fast", rng="seeded", refine_margin=margin)
        ta = rf.trace(ctx, dev, e, 0, 2048, per_test=True, **kw)
        tb = rf.trace(ctx, dev, e, 0, 2048, per_test=stacked, **kw)
        assert np.array_equal(ta, tb)
    dev.close()


def _tilted(sc, deg=23.0, axis=(0.3, 0.8, 0.52)):
    """The scene rotated about the origin: world -> object maps become invm * R^T
    and object -> world maps R * m, so coaxial bands keep equal rows apart from
    the z offset, now with x and y terms in every row (non-axial stacks)."""
    a = np.asarray(axis, dtype=np.float64)
    a /= np.linalg.norm(a)
    th = np.radians(deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    Rm = (np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K).astype(np.float32)
    p = sc.prims.copy()
    for k in range(len(p)):
        inv = p["invm"][k].reshape(3, 4)
        m = p["m"][k].reshape(3, 4)
        inv2 = np.empty_like(inv)
        inv2[:, :3] = inv[:, :3] @ Rm.T
        inv2[:, 3] = inv[:, 3]
        m2 = np.empty_like(m)
        m2[:, :3] = Rm @ m[:, :3]
        m2[:, 3] = Rm @ m[:, 3]
        p["invm"][k] = inv2.reshape(12)
        p["m"][k] = m2.reshape(12)
    return scenes.SceneData(p, sc.bvs, sc.globalRayDensity, sc.useBoundingVolumes)


@pytest.mark.parametrize("stacked", STACKED)
@pytest.mark.parametrize("margin", [0.0, 0.25])
def test_stacks_tilted(ctx, margin, stacked):
    sc = _tilted(scenes.parse_xml(scenes.furnace_xml(wall_bands=40, preform_bands=12)))
    inv = sc.prims["invm"]
    assert (inv[:40, 8] != 0).all() and (inv[:40, :11] == inv[0, :11]).all()
    a, b = _pair(ctx, sc, stacked, rays_per_emitter=20000, form="persistent", refine_margin=margin)
    _same(a, b)


# ---------------------------------------------------------------------------
# The binned bounding-surface kernel scans each surface's children with the
# same runs (the wall surface's children are one band stack).

@pytest.mark.parametrize("stacked", STACKED)
@pytest.mark.parametrize("margin", [0.0, 0.25])
@pytest.mark.parametrize("form", ["persistent", "per_emitter"])
def test_stacks_bv_binned(ctx, margin, form, stacked):
    sc = scenes.parse_xml(scenes.furnace_xml(with_bv=True, wall_bands=60, preform_bands=20))
    dev = rf.DeviceScene(ctx, sc.prims, sc.bvs)
    kw = dict(math="fast", rng="seeded", use_bvs=True, rays_per_emitter=1 << 14, form=form, refine_margin=margin)
    a = rf.solve(ctx, dev, per_test=True, **kw)
    b = rf.solve(ctx, dev, per_test=stacked, **kw)
    _same(a, b)
    assert a.timing["tests_brute_force"] == b.timing["tests_brute_force"]
    # the search looks at a few bands; the walk at every band plus a quadric
    assert (0 < b.timing["tests"] < a.timing["tests"]) if stacked == 0 else (b.timing["tests"] > a.timing["tests"])
    assert 0 < b.timing["flops"] < a.timing["flops"]
    dev.close()


def test_stacks_bv_overlapping_children(ctx):
    sc = scenes.parse_xml(scenes.furnace_xml(with_bv=True, wall_bands=60, preform_bands=20))
    bvs = sc.bvs.copy()
    bvs["objEndIndex"][2] = bvs["objStartIndex"][2]                      # no children
    bvs["objStartIndex"][1] = max(0, int(bvs["objStartIndex"][1]) - 5)   # overlaps the wall's range
    dev = rf.DeviceScene(ctx, sc.prims, bvs)
    kw = dict(math="fast", rng="seeded", use_bvs=True, rays_per_emitter=1 << 14, form="persistent")
    a = rf.solve(ctx, dev, per_test=True, **kw)
    for stacked in STACKED:
        _same(a, rf.solve(ctx, dev, per_test=stacked, **kw))
    dev.close()


# ---------------------------------------------------------------------------
# Compat (reference arithmetic) and FP64: band stacks over the caller's order,
# checked bit for bit against the CPU oracle's per-test restatement of
# findObject in the same arithmetic.

@pytest.mark.parametrize("stacked", STACKED)
@pytest.mark.parametrize("math", ["compat", "fp64"])
@pytest.mark.parametrize("scene", ["edge", "tilted", "furnace"])
def test_compat_stacks_match_oracle(ctx, oracle, scene, math, stacked):
    if scene == "edge":
        sc, n_wi = _edge_scene(), 2048
    elif scene == "tilted":
        sc, n_wi = _tilted(scenes.parse_xml(scenes.furnace_xml(wall_bands=24, preform_bands=8))), 256
    else:
        sc, n_wi = scenes.parse_xml(scenes.furnace_xml(wall_bands=60, preform_bands=20)), 128
    dev = rf.DeviceScene(ctx, sc.prims, None)
    got = rf.solve(ctx, dev, math=math, rays_per_emitter=4 * n_wi, per_test=stacked)
    want, _ = oracle.solve(sc.prims, None, n_wi=n_wi, mode=math)
    diff = np.argwhere(got.hits != want)
    assert diff.size == 0, f"{len(diff)} cells differ, first {diff[:5].tolist()}"
    assert got.timing["flops"] > 0
    dev.close()


@pytest.mark.parametrize("math", ["compat", "fp64"])
@pytest.mark.parametrize("form", ["persistent", "per_emitter", "reference"])
def test_compat_stacks_equal_per_test(ctx, form, math):
    geoms = scenes.sweep(3, seed=0x52465331)
    for g in geoms:
        dev = rf.DeviceScene(ctx, g.prims, None)
        kw = dict(math=math, rng="seeded", rays_per_emitter=1 << 13, form=form)
        a = rf.solve(ctx, dev, per_test=True, **kw)
        for stacked in STACKED:
            b = rf.solve(ctx, dev, per_test=stacked, **kw)
            assert np.array_equal(a.hits, b.hits)
            assert b.timing["flops"] < a.timing["flops"]
        dev.close()


@pytest.mark.parametrize("stacked", STACKED)
@pytest.mark.parametrize("math", ["compat", "fp64"])
def test_compat_stacks_traces(ctx, oracle, math, stacked):
    sc = _edge_scene()
    dev = rf.DeviceScene(ctx, sc.prims, None)
    for e in (0, 5, 8, 12, 16, 19):
        a = rf.trace(ctx, dev, e, 100, 1500, math=math, per_test=stacked)
        b = oracle.trace(sc.prims, None, e, 100, 1500, mode=math)
        assert np.array_equal(a, b), (e, np.argwhere(a != b)[:5])
    dev.close()


@pytest.mark.parametrize("stacked", STACKED)
@pytest.mark.parametrize("math", ["compat", "fp64"])
def test_compat_stacks_bv_match_oracle(ctx, oracle, math, stacked):
    """runBVGPU semantics: each surface's child range is scanned with the band
    stacks of the caller's order cut to that range.  One surface's range is
    moved to start five bands inside the wall's stack."""
    sc = scenes.parse_xml(scenes.furnace_xml(with_bv=True, wall_bands=60, preform_bands=20))
    bvs = sc.bvs.copy()
    bvs["objStartIndex"][1] = max(0, int(bvs["objStartIndex"][1]) - 5)
    dev = rf.DeviceScene(ctx, sc.prims, bvs)
    n_wi = 128
    got = rf.solve(ctx, dev, math=math, use_bvs=True, rays_per_emitter=4 * n_wi, per_test=stacked)
    want, _ = oracle.solve(sc.prims, bvs, n_wi=n_wi, mode=math, use_bvs=True)
    diff = np.argwhere(got.hits != want)
    assert diff.size == 0, f"{len(diff)} cells differ, first {diff[:5].tolist()}"
    a = rf.solve(ctx, dev, math=math, use_bvs=True, rays_per_emitter=4 * n_wi, per_test=True)
    assert np.array_equal(a.hits, got.hits) and a.timing["tests_brute_force"] == got.timing["tests_brute_force"]
    # the search looks at a few bands; the walk at every band plus a quadric
    assert (0 < got.timing["tests"] <= a.timing["tests"]) if stacked == 0 else (got.timing["tests"] >= a.timing["tests"])
    dev.close()


def _uneven_scene(seed=7):
    """Stacks whose bands are in z order but unevenly spaced (random gaps and
    overlaps), so the even-spacing guess often misses and the binary search
    decides: 80 wall bands of height 0.05 listed bottom up, 30 inner bands of
    height 0.1 listed top down, closed by a disc at each end."""
    rng = np.random.default_rng(seed)
    z_wall = np.sort(rng.uniform(0.0, 3.95, 80))
    z_in = np.sort(rng.uniform(0.5, 3.0, 30))[::-1]
    parts = [_band(z, z + 0.05) for z in z_wall]
    parts += [_band(z, z + 0.1, r=0.2) for z in z_in]
    parts += [scenes._prim("disc", scale=(0.5, 0.5, 1)),
              scenes._prim("disc", scale=(0.5, 0.5, 1), rotate=scenes.FLIP_X, translate=(0, 0, 4.0))]
    return scenes.parse_xml(scenes._doc("unevenly spaced band stacks", "".join(parts)))


@pytest.mark.parametrize("margin", [0.0, 0.25])
@pytest.mark.parametrize("rpt", [1, 2])
def test_stacks_uneven_spacing(ctx, margin, rpt):
    sc = _uneven_scene()
    a, b = _pair(ctx, sc, 0, rays_per_emitter=1 << 14, form="persistent", rays_per_thread=rpt,
                 refine_margin=margin)
    _same(a, b)
    assert a.hits[:, :80].sum() > 0 and a.hits[:, 80:110].sum() > 0


@pytest.mark.parametrize("math", ["compat", "fp64"])
def test_compat_stacks_uneven_spacing(ctx, oracle, math):
    sc = _uneven_scene()
    dev = rf.DeviceScene(ctx, sc.prims, None)
    n_wi = 256
    got = rf.solve(ctx, dev, math=math, rays_per_emitter=4 * n_wi)
    want, _ = oracle.solve(sc.prims, None, n_wi=n_wi, mode=math)
    diff = np.argwhere(got.hits != want)
    assert diff.size == 0, f"{len(diff)} cells differ, first {diff[:5].tolist()}"
    dev.close()


@pytest.mark.parametrize("math", ["fast", "compat", "fp64"])
def test_executed_evaluation_counts(ctx, math):
    """rf_timing.tests counts what runs: every test in the per-test form
    (rays x n); in the walk form one quadric per (ray, stack) plus one z-range
    check per band (exactly rays x (n outside stacks + sum of (1 + bands)));
    in the search form fewer.  tests_brute_force is rays x n in all three."""
    wall, pre = 40, 12
    sc = scenes.parse_xml(scenes.furnace_xml(wall_bands=wall, preform_bands=pre))
    dev = rf.DeviceScene(ctx, sc.prims, None)
    rays = 1 << 14
    kw = dict(math=math, rng="seeded", rays_per_emitter=rays, form="persistent", rays_per_thread=4 if math == "fast"
              else 0)
    full = rf.solve(ctx, dev, per_test=1, **kw).timing
    walk = rf.solve(ctx, dev, per_test=2, **kw).timing
    search = rf.solve(ctx, dev, per_test=0, **{**kw, "rays_per_thread": 2 if math == "fast" else 0}).timing
    n_rays = rays * sc.n                       # every emitter
    assert full["tests"] == full["tests_brute_force"] == n_rays * sc.n
    per_ray = (sc.n - wall - pre) + (1 + wall) + (1 + pre)
    assert walk["tests"] == n_rays * per_ray and walk["tests_brute_force"] == n_rays * sc.n
    assert search["tests_brute_force"] == n_rays * sc.n
    assert 0 < search["tests"] < walk["tests"]
    print(f"[evals] {math}: evaluations per ray: per test {sc.n}, walk {per_ray}, "
          f"search {search['tests'] / (rays * sc.n):.2f}", flush=True)
    dev.close()


@pytest.mark.parametrize("rpt", [1, 2])
def test_band_tile_and_counters_keep_tallies(ctx, rpt):
    """The band search reading its records from an LDS copy, and without its
    evaluation counters, gives the same tallies and re-solves; the reported
    shape says which ran; uncounted solves report tests / flops as -1."""
    from rayfactorcl_amd import _lib as L
    geoms = scenes.sweep(5, seed=0x52465331) + [_uneven_scene()]
    dev = [rf.DeviceScene(ctx, g.prims, None) for g in geoms]
    kw = dict(math="fast", rng="seeded", rays_per_emitter=1 << 14, form="persistent", rays_per_thread=rpt)
    ref, tr = rf.batch_solve(ctx, dev, rf.make_opts(band_tile=-1, **kw))
    assert tr["shape_flags"] & L.RF_SHAPE_SEARCHED and not tr["shape_flags"] & L.RF_SHAPE_BAND_TILE
    # explicit tiles: the session's context may have been tuned (tile 0 = its choice)
    for tile, count in [(8192, True), (8192, False), (-1, False), (4096, True)]:
        got, t = rf.batch_solve(ctx, dev, rf.make_opts(band_tile=tile, count_tests=count, **kw))
        for a, b in zip(ref, got):
            assert np.array_equal(a.hits, b.hits), (tile, count)
        assert t["refined"] == tr["refined"]
        assert bool(t["shape_flags"] & L.RF_SHAPE_BAND_TILE) == (tile != -1)
        assert bool(t["shape_flags"] & L.RF_SHAPE_COUNTED) == count
        if count:
            assert t["tests"] == tr["tests"] and t["flops"] == tr["flops"]
        else:
            assert t["tests"] == -1 and t["flops"] == -1
    # a tile smaller than a scene's cylinder segment keeps the records in global memory
    _, t = rf.batch_solve(ctx, dev, rf.make_opts(band_tile=8, **kw))
    assert not t["shape_flags"] & L.RF_SHAPE_BAND_TILE
    for d in dev:
        d.close()


def _shift_z(prims, dz):
    """The scene moved by dz along world z: m's z translation and invm's
    offsets, rounded to FP32 like any loaded scene."""
    p = prims.copy()
    m = p["m"].astype(np.float64)
    iv = p["invm"].astype(np.float64)
    m[:, 11] += dz
    for r in range(3):
        iv[:, 4 * r + 3] -= iv[:, 4 * r + 2] * dz
    p["m"] = m.astype(np.float32)
    p["invm"] = iv.astype(np.float32)
    return p


@pytest.mark.parametrize("dz, bands", [(0.0, 500), (900.0, 180), (-2.5e4, 60)])
def test_analytic_positions_equal_per_test(ctx, dz, bands):
    """Analytic band positions: each root of an evenly spaced axial stack is
    placed in FP64 and taken without reading a band record only where the
    margin (tau, the offsets' deviation from their progression, and the FP32
    rounding of the per-band z) proves every per-band test agrees.  Far from
    the origin the offsets are large, their rounding makes the progression
    uneven and the margin wide, so more slots fall back to the probe search;
    either way the tallies, the re-solve set and the traces are the per-test
    form's."""
    sc = scenes.parse_xml(scenes.furnace_xml(wall_bands=bands, preform_bands=max(12, bands // 3)))
    prims = _shift_z(sc.prims, dz)
    dev = rf.DeviceScene(ctx, prims, None)
    kw = dict(math="fast", rng="seeded", rays_per_emitter=1 << 15, form="persistent", rays_per_thread=2,
              block_size=256)
    a = rf.solve(ctx, dev, per_test=True, **kw)
    b = rf.solve(ctx, dev, per_test=0, count_tests=True, **kw)
    _same(a, b)
    for e in (0, bands // 2, len(prims) - 1):
        ta = rf.trace(ctx, dev, e, 0, 4096, math="fast", rng="seeded", per_test=True)
        tb = rf.trace(ctx, dev, e, 0, 4096, math="fast", rng="seeded", per_test=0)
        assert np.array_equal(ta, tb), e
    # fewer evaluations than the per-test form's n per ray (far from the
    # origin more slots go through the search's probes)
    assert 0 < b.timing["tests"] < 0.25 * a.timing["tests"]
    print(f"[analytic] dz={dz} bands={bands}: {b.timing['tests'] / b.rays.sum():.2f} evaluations per ray "
          f"(per test {a.timing['tests'] / a.rays.sum():.0f})", flush=True)
    dev.close()


@pytest.mark.parametrize("dz, bands", [(0.0, 500), (-2.5e4, 60)])
def test_compat_analytic_positions_equal_per_test(ctx, oracle, dz, bands):
    """The compat band stacks place each root analytically too (their range
    tests z1 in [0, 1] and z2 in (0, 1) decided with a derived margin):
    searched, per test and the CPU oracle's per-test scan agree bit for bit,
    also far from the origin where the offsets' rounding widens the margin."""
    sc = scenes.parse_xml(scenes.furnace_xml(wall_bands=bands, preform_bands=max(12, bands // 3)))
    prims = _shift_z(sc.prims, dz)
    dev = rf.DeviceScene(ctx, prims, None)
    kw = dict(math="compat", rng="seeded", rays_per_emitter=1 << 14, form="persistent")
    a = rf.solve(ctx, dev, per_test=True, **kw)
    b = rf.solve(ctx, dev, per_test=0, **kw)
    assert np.array_equal(a.hits, b.hits)
    for e in (0, bands // 2, len(prims) - 1):
        got = rf.trace(ctx, dev, e, 0, 2048, math="compat", rng="seeded")
        want = oracle.trace(prims, None, e, 0, 2048, mode="compat", rng="seeded")
        assert np.array_equal(got, want), e
    dev.close()
