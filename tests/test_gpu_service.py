"""Streaming service (resident scenes, HBM-footprint admission, batched
flushes) and the autotuner."""
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
    r1 = configs.config1(ctx, quick=True)
    assert all(c["within_4sigma_plus_edge"] for c in r1["fast"]["closed_forms"])
    assert all(c["within_4sigma_plus_edge"] for c in r1["compat"]["closed_forms"])
    assert r1["fast_vs_fp64_max_dvf"] <= 1e-4 * 10     # 1e6 rays: binomial noise of differing rays
    r5 = configs.config5(ctx, quick=True, n_req=6)
    assert r5["requests"] == 6 and r5["latency_p99_s"] >= r5["latency_p50_s"] > 0
