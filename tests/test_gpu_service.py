"""Streaming service (resident scenes, HBM-footprint admission, batched
flushes) and the autotuner."""
import math

import numpy as np
import pytest

import rayfactorcl_amd as rf
from rayfactorcl_amd import _lib, scenes

pytestmark = pytest.mark.gpu


def test_stream_matches_direct_solves(ctx):
    st = rf.Stream(ctx, 1 << 30)
    reqs = []
    for k, (n, rays) in enumerate([(30, 5000), (200, 2000), (57, 5000), (120, 800)]):
        sc = scenes.parse_xml(scenes.soup_xml(n, seed=100 + k))
        o = rf.make_opts(math="fast", rng="seeded", rays_per_emitter=rays, seed=3, geometry=k)
        reqs.append((st.submit(sc.prims, o), sc, rays, k))
    assert st.resident_bytes() > 0
    assert st.flush() == 4
    for rid, sc, rays, k in reqs:
        h, r, lat = st.collect(rid)
        assert lat >= 0 and list(r) == [4 * ((rays + 3) // 4)] * sc.n
        dev = rf.DeviceScene(ctx, sc.prims, None)
        one = rf.solve(ctx, dev, math="fast", rng="seeded", rays_per_emitter=rays, seed=3, geometry=k)
        assert np.array_equal(h, one.hits)
    assert st.resident_bytes() == 0
    st.close()


def test_stream_admission(ctx):
    big = scenes.parse_xml(scenes.soup_xml(400, seed=1))
    limit = rf.scene_footprint(400) + 10
    st = rf.Stream(ctx, limit)
    o = rf.make_opts(math="fast", rng="seeded", rays_per_emitter=1000)
    rid = st.submit(big.prims, o)
    with pytest.raises(rf.RayFactorError) as e:
        st.submit(big.prims, o)
    assert e.value.code == _lib.RF_ERR_CAPACITY
    st.flush()
    st.collect(rid)
    st.submit(big.prims, o)           # admitted again once the first left
    st.close()


def test_stream_mixed_options_each_equal_their_own_solve(ctx):
    """Requests that differ only in the FP64 re-solve margin or the transform
    unit are never solved with another request's options."""
    sc = scenes.parse_xml(scenes.soup_xml(300, seed=11))
    st = rf.Stream(ctx, 1 << 30)
    variants = [dict(), dict(refine_margin=0.25), dict(transform="mfma"), dict(refine=False),
                dict(refine_margin=0.25, transform="mfma"), dict(seed=6), dict(math="fp64")]
    base = dict(math="fast", rng="seeded", rays_per_emitter=1 << 16, seed=5)
    ids = [st.submit(sc.prims, rf.make_opts(**{**base, **v})) for v in variants]
    assert st.flush() == len(variants)
    dev = rf.DeviceScene(ctx, sc.prims, None)
    got = [st.collect(rid)[0] for rid in ids]
    for v, h in zip(variants, got):
        assert np.array_equal(h, rf.solve(ctx, dev, **{**base, **v}).hits), v
    # another seed or arithmetic form certainly changes the tallies; on 2e7 soup
    # rays so does a margin, transform or re-solve change (a few rays each)
    assert not np.array_equal(got[0], got[5]) and not np.array_equal(got[0], got[6])
    assert not np.array_equal(got[0], got[2]) or not np.array_equal(got[0], got[3])
    print(f"[stream] tallies differ from the default: margin {not np.array_equal(got[0], got[1])}, "
          f"mfma {not np.array_equal(got[0], got[2])}, unrefined {not np.array_equal(got[0], got[3])}", flush=True)
    st.close()


def test_stream_failed_flush_keeps_requests(ctx):
    """A flush whose launch fails leaves every uncompleted request queued and
    its footprint accounted; the next flush completes them all."""
    st = rf.Stream(ctx, 1 << 30)
    o1 = rf.make_opts(math="fast", rng="seeded", rays_per_emitter=4000, seed=1)
    o2 = rf.make_opts(math="compat", rng="seeded", rays_per_emitter=4000, seed=1)
    scs = [scenes.parse_xml(scenes.soup_xml(n, seed=20 + n)) for n in (20, 40, 60)]
    ids = [st.submit(scs[0].prims, o1), st.submit(scs[1].prims, o2), st.submit(scs[2].prims, o1)]
    resident = st.resident_bytes()
    rf.fail_solve(ctx, 2)                       # the first launch runs, the second fails
    with pytest.raises(rf.RayFactorError) as e:
        st.flush()
    assert e.value.code == _lib.RF_ERR_HIP and "injected" in str(e.value)
    assert st.resident_bytes() == resident
    assert st.queued_cost() > 0
    assert st.flush() == 1                      # the request left over from the failed launch
    assert st.queued_cost() == 0
    for rid, sc, o in zip(ids, scs, (o1, o2, o1)):
        h, _, _ = st.collect(rid)
        dev = rf.DeviceScene(ctx, sc.prims, None)
        assert np.array_equal(h, rf.solve(ctx, dev, o).hits)
    assert st.resident_bytes() == 0
    st.close()


def test_stream_cost_admission_and_packing(ctx):
    """Cost admission rejects a submit past the queued-cost limit; packing
    runs the cheapest requests first, so they complete no later than the
    costly one."""
    sc_small = scenes.parse_xml(scenes.soup_xml(20, seed=1))
    sc_big = scenes.parse_xml(scenes.soup_xml(300, seed=2))
    o_small = rf.make_opts(math="fast", rng="seeded", rays_per_emitter=4000)
    o_big = rf.make_opts(math="fast", rng="seeded", rays_per_emitter=40000)
    small_cost = 4000 * 20 * 20
    # 300 primitives outside band stacks: the hierarchy's scan-equivalent
    # cost, 6 sqrt(300) tests per ray (20 get no hierarchy)
    big_cost = 40000 * 300 * 6.0 * math.sqrt(300)
    st = rf.Stream(ctx, 1 << 30)
    st.set_cost_limits(max_queued_tests=big_cost + 3 * small_cost, max_launch_tests=4 * small_cost)
    big = st.submit(sc_big.prims, o_big)
    assert st.queued_cost() == pytest.approx(big_cost)
    small = [st.submit(sc_small.prims, o_small) for _ in range(3)]
    with pytest.raises(rf.RayFactorError) as e:
        st.submit(sc_small.prims, o_small)
    assert e.value.code == _lib.RF_ERR_CAPACITY
    assert st.flush() == 4
    lat_big = st.collect(big)[2]
    lat_small = [st.collect(r)[2] for r in small]
    assert max(lat_small) <= lat_big
    st.submit(sc_small.prims, o_small)          # admitted again once the queue drained
    st.close()


def test_stream_cost_of_the_fp64_walk(ctx):
    """An FP64 request over a scene whose hierarchy holds every primitive is
    weighed at the walk's scan-equivalent cost, 6 sqrt(m) tests per ray; with
    the hierarchy switched off, at the full scan's m."""
    sc = scenes.parse_xml(scenes.soup_xml(300, seed=2))
    st = rf.Stream(ctx, 1 << 30)
    walk = st.submit(sc.prims, rf.make_opts(math="fp64", rng="seeded", rays_per_emitter=1000))
    walk_cost = 1000 * 300 * 6.0 * math.sqrt(300)
    assert st.queued_cost() == pytest.approx(walk_cost)
    scan = st.submit(sc.prims, rf.make_opts(math="fp64", rng="seeded", rays_per_emitter=1000,
                                            hierarchy=False))
    assert st.queued_cost() == pytest.approx(walk_cost + 1000 * 300 * 300)
    assert st.flush() == 2
    assert np.array_equal(st.collect(walk)[0], st.collect(scan)[0])
    st.close()


def test_stream_chunks_a_large_request(ctx):
    """A request costlier than the launch limit runs in chunks of emitters,
    one per step; its tallies are rf_solve's, and a small request submitted
    between two chunks completes before the large one."""
    big = scenes.parse_xml(scenes.soup_xml(120, seed=31))
    small = scenes.parse_xml(scenes.soup_xml(15, seed=32))
    ob = rf.make_opts(math="fast", rng="seeded", rays_per_emitter=20000, seed=7)
    os_ = rf.make_opts(math="fast", rng="seeded", rays_per_emitter=4000, seed=7)
    st = rf.Stream(ctx, 1 << 30)
    cost_big = 20000 * 120 * 6.0 * math.sqrt(120)     # the hierarchy's scan-equivalent cost
    st.set_cost_limits(0.0, cost_big / 6)              # about 20 emitters per chunk
    rb = st.submit(big.prims, ob)
    steps, done = 0, 0
    done += st.step()                                  # first chunk of the large request
    steps += 1
    assert done == 0 and 0 < st.queued_cost() < cost_big
    rs = st.submit(small.prims, os_)                   # cheaper than what is left: runs next
    done += st.step()
    steps += 1
    assert done == 1
    hs, _, lat_s = st.collect(rs)
    while done < 2:
        done += st.step()
        steps += 1
    hb, rays_b, lat_b = st.collect(rb)
    assert steps >= 7 and lat_s < lat_b
    assert st.queued_cost() == 0 and st.resident_bytes() == 0
    for sc, o, h in ((big, ob, hb), (small, os_, hs)):
        dev = rf.DeviceScene(ctx, sc.prims, None)
        ref = rf.solve(ctx, dev, o)
        assert np.array_equal(h, ref.hits)
        dev.close()
    assert list(rays_b) == [20000] * big.n
    st.close()


def test_autotune_keeps_results(ctx):
    """Tuning picks a launch shape and a transform; the VALU kernel's tallies
    do not depend on the shape, and AUTO afterwards equals the chosen
    transform pinned."""
    geoms = scenes.sweep(3, seed=4)
    dev = [rf.DeviceScene(ctx, g.prims, None) for g in geoms]
    kw = dict(math="fast", rng="seeded", rays_per_emitter=2000, form="persistent")
    before = {tr: rf.batch_solve(ctx, dev, transform=tr, **kw)[0] for tr in ("valu", "mfma")}
    rpt, blk = rf.autotune(ctx, dev, math="fast", rng="seeded", rays_per_emitter=2000)
    assert rpt in (0, 2, 4, 8) and blk in (128, 256)
    after_valu, _ = rf.batch_solve(ctx, dev, transform="valu", **kw)
    after_auto, _ = rf.batch_solve(ctx, dev, **kw)
    want = before["mfma" if rpt == 0 else "valu"]
    for a, b, c, w in zip(before["valu"], after_valu, after_auto, want):
        assert np.array_equal(a.hits, b.hits)
        assert np.array_equal(c.hits, w.hits)


def test_configs_report_quick(ctx):
    """The per-configuration report runs end to end (reduced budgets)."""
    from rayfactorcl_amd import configs
    r1 = configs.config1(ctx, quick=True, cpu_seconds=0)
    assert all(c["within_3se"] for c in r1["fast"]["closed_forms"])
    assert all(c["within_3se"] for c in r1["compat"]["closed_forms"])
    assert r1["fast_vs_fp64_max_dvf"] <= 1e-4 * 10     # 1e6 rays: binomial noise of differing rays
    r5 = configs.config5(ctx, quick=True, n_req=6)
    assert r5["requests"] == 6 and r5["latency_p99_s"] >= r5["latency_p50_s"] > 0


def test_closure_reciprocity_report(ctx):
    """Report of a solved closed-cylinder configuration: the enclosure rows
    close to 1 (less the seam losses of the EPS guard), the open coaxial pair sees its
    partner at the disc-to-disc view factor, and A_i F_ij matches A_j F_ji
    within the Monte Carlo error."""
    sc = scenes.parse_xml(scenes.config1_xml())
    dev = rf.DeviceScene(ctx, sc.prims, None)
    rays = 4_000_000
    r = rf.solve(ctx, dev, math="fast", rng="seeded", rays_per_emitter=rays)
    rep = rf.vf_report(r.vf, rf.areas(sc.prims), min_product=1e-2)
    dev.close()
    c = rep["closure"]
    # rays leaving within EPS of a seam meet the neighbouring surface at
    # t <= EPS and are not counted (the reference's self-hit guard): each row
    # of the enclosure closes to 1 less its rims' seam losses.  The standard
    # error of a miss fraction m is at most 2 sqrt(m / rays): the 4 rays of a
    # work item share a random block, which at worst doubles a binomial sigma
    from closed_forms import REMAINDER, seam_loss
    lost = np.array([2 * seam_loss(2 * np.pi, 4 * np.pi, np.pi / 2), seam_loss(2 * np.pi, np.pi, np.pi / 2),
                     seam_loss(2 * np.pi, np.pi, np.pi / 2)])
    assert np.all(np.abs(c[:3] - (1.0 - lost)) <= 3 * 2 * np.sqrt(lost / rays) + REMAINDER), (c[:3], 1 - lost)
    f_pp = 0.5 * (3.0 - 5.0 ** 0.5)
    sig = 2.0 * (f_pp * (1 - f_pp) / rays) ** 0.5
    # the pair also sees the enclosure's outside at grazing angles (< 1e-2)
    for i, j in ((3, 4), (4, 3)):
        assert abs(r.vf[i, j] - f_pp) < 5 * sig
        assert r.vf[i, j] <= c[i] < r.vf[i, j] + 1e-2
    assert rep["pairs_compared"] == 4
    assert rep["max_recip_rel"] < 8e-3


def test_cli_output_and_report(ctx, tmp_path):
    """The drop-in program writes the same output.txt as a library solve with
    the same options and prints the closure / reciprocity report."""
    import subprocess
    from rayfactorcl_amd.build import CLI
    xml = tmp_path / "input.xml"
    xml.write_text(scenes.config1_xml())
    out = tmp_path / "output.txt"
    rays = 1_000_000
    p = subprocess.run([str(CLI), str(xml), "--math", "fast", "--rays", str(rays), "--out", str(out), "--report"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "Closure : min" in p.stdout and "Reciprocity : 4 pairs" in p.stdout
    sc = scenes.parse_xml(scenes.config1_xml())
    dev = rf.DeviceScene(ctx, sc.prims, None)
    r = rf.solve(ctx, dev, math="fast", rays_per_emitter=rays)
    dev.close()
    rf.write_output(str(tmp_path / "expected.txt"), r.hits, r.rays.astype(np.float64))
    assert out.read_text() == (tmp_path / "expected.txt").read_text()
