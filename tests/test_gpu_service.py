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
    # t <= EPS and are not counted (the reference's self-hit guard): an edge
    # loss of a few 1e-4, bounded as the closed-form checks bound it
    assert np.all(c[:3] <= 1.0) and np.all(c[:3] >= 1.0 - 1.5e-3)
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
