"""Philox4x32-10 and the u01 conversion of the oracle, pinned by the published
Random123 known-answer vectors and by the reference's own philox.h (compiled
from its sources into oracle/_ref; outputs in tests/golden/philox_vectors.json)."""
import json

import numpy as np
import pytest

from conftest import GOLDEN

# Published Random123 KATs for philox4x32_10 (kat_vectors of Random123 1.06)
PUBLISHED = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", PUBLISHED)
def test_philox_published_kat(oracle, ctr, key, out):
    assert list(oracle.philox(ctr, key)) == out


def test_philox_matches_reference_header(oracle):
    vecs = json.loads((GOLDEN / "philox_vectors.json").read_text())
    assert len(vecs) > 60
    for v in vecs:
        assert list(oracle.philox(v["ctr"], v["key"])) == v["out"], v


def test_reference_stream_first_items(oracle):
    # RayFactorGPU.cl:486-499 for gid 0 (values measured from the reference header)
    u = oracle.workitem_u32("reference", 0, 0, 0, 0)
    assert list(u) == [0x0A9E1FA3, 0xBB160856, 0x266DF5BA, 0x11BA516A]
    f = [oracle.u01(int(x)) for x in u]
    assert f == pytest.approx([0.0414752364, 0.73080492, 0.150115311, 0.0692492127], abs=1e-9)


def test_workitem_keys_follow_fixture(oracle):
    vecs = json.loads((GOLDEN / "philox_vectors.json").read_text())
    for v in vecs:
        if v["kind"] == "reference":
            got = oracle.workitem_u32("reference", 0, 0, 0, v["gid"])
        elif v["kind"] == "seeded":
            got = oracle.workitem_u32("seeded", v["seed"], v["geometry"], v["emitter"], v["wi"])
        else:
            continue
        assert list(got) == v["out"], v


def test_reference_stream_ignores_emitter(oracle):
    # Q3: the same key/counter for every emitter
    a = oracle.workitem_u32("reference", 5, 9, 0, 77)
    b = oracle.workitem_u32("reference", 6, 1, 3, 77)
    assert list(a) == list(b)
    c = oracle.workitem_u32("seeded", 5, 9, 0, 77)
    d = oracle.workitem_u32("seeded", 5, 9, 3, 77)
    assert list(c) != list(d)


@pytest.mark.parametrize("u", [0, 1, 255, 256, 0x7FFFFFFF, 0x80000000, 0xFFFFFF00, 0xFFFFFFFF, 0x12345678])
def test_u01_closed_form(oracle, u):
    # (u >> 8) * 2^-24 is exact in float
    assert oracle.u01(u) == np.float32((u >> 8) * 2.0 ** -24)
    assert 0.0 <= oracle.u01(u) < 1.0
