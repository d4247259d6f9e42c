"""Synthetic workloads: sizes and structure of the generated scenes."""
import numpy as np

import rayfactorcl_amd as rf
from rayfactorcl_amd import scenes


def test_furnace_config():
    sc = scenes.parse_xml(scenes.furnace_xml())
    assert sc.n == 244
    t = sc.prims["type"]
    assert (t == rf.CYLINDER).sum() == 240 and (t == rf.FRUSTUM).sum() == 1
    assert (t == rf.ANNULUS).sum() == 2 and (t == rf.DISC).sum() == 1
    assert (sc.prims["util2"][:180] == -1).all() and (sc.prims["util2"][180:240] == 1).all()
    bv = scenes.parse_xml(scenes.furnace_xml(with_bv=True))
    assert bv.useBoundingVolumes and bv.n == 244 and len(bv.bvs) == 6
    assert bv.bvs["objEndIndex"][0] - bv.bvs["objStartIndex"][0] == 180
    # same primitives in both forms
    assert bv.prims.tobytes() == sc.prims.tobytes()


def test_sweep_ranges():
    ps = scenes.sweep_params(256)
    ns = []
    for p in ps:
        assert 20 <= p["wall_bands"] <= 500 and 10 <= p["preform_bands"] <= 100
        assert 0.3 <= p["wall_r"] <= 0.8 and 1 <= p["height"] <= 4 and 0.05 <= p["preform_r"] <= 0.2
        ns.append(p["wall_bands"] + p["preform_bands"] + 3 + (p["neck_deg"] is not None))
    assert 33 <= min(ns) and max(ns) <= 605
    g = scenes.sweep(4)
    assert [s.n for s in g] == ns[:4]


def test_soup():
    sc = scenes.parse_xml(scenes.soup_xml(200, seed=5))
    assert sc.n == 200
    counts = np.bincount(sc.prims["type"], minlength=6)
    assert counts[rf.RECTANGLE] == 0 and counts[[0, 1, 2, 3, 5]].min() > 20


def test_config1_xml_matches_fixture():
    from conftest import GOLDEN
    a = scenes.parse_xml(scenes.config1_xml())
    b = rf.read_rfscene(GOLDEN / "scenes" / "config1.rfscene")
    assert a.prims.tobytes() == b.prims.tobytes()
