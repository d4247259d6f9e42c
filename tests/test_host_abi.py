"""Host-side parts of the C ABI (no GPU compute): the library loads and exports
every declared entry point, struct layouts, launch geometry, the scene loader
(byte-identical to the reference loader's output), the output writer and
error paths."""
import re
from pathlib import Path

import numpy as np
import pytest

import rayfactorcl_amd as rf
from rayfactorcl_amd import _lib
from conftest import GOLDEN, ROOT, SCENE_NAMES


def declared_functions():
    text = (ROOT / "include" / "rayfactor.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    L = rf.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_struct_layouts():
    assert rf.PRIM_DTYPE.itemsize == 112 and rf.BV_DTYPE.itemsize == 72
    assert rf.PRIM_DTYPE.fields["type"][1] == 96 and rf.PRIM_DTYPE.fields["util2"][1] == 108
    assert rf.BV_DTYPE.fields["objEndIndex"][1] == 60 and rf.BV_DTYPE.fields["util2"][1] == 68
    assert rf.lib().rf_abi_version() == 3


# work items per emitter of the sample scenes at globalRayDensity = 1e6
# (main.cpp:172-173 with primitive 0's rayDensity, Q4, and nearest-128 rounding, Q5)
KNOWN = [("input-C14", 999936), ("input-C47", 749952), ("input-C109", 749952), ("input-C52", 499968),
         ("input-C77", 2999936), ("input-C122", 2999936), ("input-furnace", 50249984), ("input-W100", 50249984)]


@pytest.mark.parametrize("scene,expect", KNOWN)
def test_launch_geometry_known_answers(golden_scene, oracle, scene, expect):
    sc = golden_scene(scene)
    rd = int(sc.prims[0]["rayDensity"])
    assert rf.nthreads(rd, int(sc.globalRayDensity)) == expect
    assert oracle.nthreads(rd, int(sc.globalRayDensity)) == expect


def test_launch_geometry_edges(oracle):
    assert rf.nthreads(1, 64 * 4) == 0            # 64 % 128 == 64: ties round down to 0
    assert rf.nthreads(1, 65 * 4) == 128
    assert rf.nthreads(1, 192 * 4) == 128         # 192 % 128 = 64 -> down
    assert rf.nthreads(1, 193 * 4) == 256
    with pytest.raises(rf.RayFactorError) as e:
        rf.nthreads(50000, 100000)                # rayDensity * n_samples overflows int32
    assert e.value.code == _lib.RF_ERR_OVERFLOW
    assert oracle.nthreads(50000, 100000) == -1


@pytest.mark.parametrize("name", SCENE_NAMES)
def test_loader_matches_reference_bytes(name):
    ref = rf.read_rfscene(GOLDEN / "scenes" / f"{name}.rfscene")
    got = rf.load_xml(GOLDEN / "xml" / f"{name}.xml")
    assert got.prims.tobytes() == ref.prims.tobytes()
    assert got.bvs.tobytes() == ref.bvs.tobytes()
    assert got.globalRayDensity == ref.globalRayDensity
    assert got.useBoundingVolumes == ref.useBoundingVolumes


def test_loader_quirks():
    c47 = rf.load_xml(GOLDEN / "xml" / "input-C47.xml")
    # flipped annulus: world normal (0, -8.74e-8, -1), squared radii in util1/util2 (Q16)
    assert c47.prims[1]["type"] == rf.ANNULUS and tuple(c47.prims[1][["util1", "util2"]]) == (1.0, 4.0)
    assert c47.prims[1]["invm"][10] == -1.0 and c47.prims[1]["invm"][9] == np.float32(-8.742278e-08)
    c109 = rf.load_xml(GOLDEN / "xml" / "input-C109.xml")
    assert c109.prims[0]["type"] == rf.FRUSTUM and c109.prims[0]["util1"] == 0.0      # Q14: cone
    assert c109.prims[0]["util2"] == -1.0                                             # Q15: inward
    uni = rf.load_xml(GOLDEN / "xml" / "input-uni.xml")                               # Q13
    assert not uni.useBoundingVolumes and len(uni.bvs) == 0 and uni.n == 1
    bv = rf.load_xml(GOLDEN / "xml" / "input-C14-BV.xml")
    assert bv.useBoundingVolumes and list(bv.bvs["objStartIndex"]) == [0, 1] and list(bv.bvs["objEndIndex"]) == [1, 2]


def test_loader_errors(tmp_path):
    with pytest.raises(rf.RayFactorError) as e:
        rf.load_xml(tmp_path / "missing.xml")
    assert e.value.code == _lib.RF_ERR_IO
    for bad in ["<project><settings>", "<project></project>", "not xml at all <",
                '<project><settings><globalRayDensity value="1"/></settings><geometry>'
                '<primitive type="annulus"><smallRadius value="1"/></primitive></geometry></project>']:
        with pytest.raises(rf.RayFactorError) as e:
            rf.parse_xml(bad)
        assert e.value.code == _lib.RF_ERR_PARSE
    # unknown primitive types are dropped (Scene.cpp:270-279)
    sc = rf.parse_xml('<project><settings><globalRayDensity value="5"/></settings><geometry>'
                      '<primitive type="triangle"><scale x="1" y="1" z="1"/></primitive>'
                      '<primitive type="disc"><scale x="2" y="2" z="1"/></primitive></geometry></project>')
    assert sc.n == 1 and sc.prims[0]["type"] == rf.DISC and sc.globalRayDensity == 5.0


def test_output_writer_format(tmp_path):
    hits = np.array([[0, 3, 17], [5, 0, 0], [123456789, 1, 2]], dtype=np.int64)
    samples = np.array([12.0, 7.0, 999936.0 * 4])
    path = tmp_path / "output.txt"
    rf.write_output(str(path), hits, samples)
    lines = path.read_text().splitlines()
    assert len(lines) == 3
    for i, line in enumerate(lines):
        assert line.endswith(",")
        vals = line[:-1].split(",")
        for j, v in enumerate(vals):
            f = np.float32(np.float32(hits[i, j]) / samples[i])
            assert v == ("%.8g" % f)
    vf = rf.normalise(hits, np.array([12, 7, 0]))
    assert vf[0, 2] == 17 / 12 and np.all(vf[2] == 0)


def test_no_device_is_a_status(monkeypatch):
    monkeypatch.setenv("HIP_VISIBLE_DEVICES", "-1")
    import subprocess
    import sys
    code = ("import rayfactorcl_amd as rf\n"
            "try:\n    rf.Context(0)\nexcept rf.RayFactorError as e:\n    print('status', e.code)\n")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=str(ROOT), timeout=120)
    assert "status -9" in out.stdout or "status -3" in out.stdout, out.stdout + out.stderr


def test_scene_footprint():
    assert rf.scene_footprint(100, 0) > 100 * 100 * 8
    assert rf.scene_footprint(-1, 0) == -1


def test_ctypes_mirrors_header_structs(tmp_path):
    """The ctypes structures match the C compiler's layout of include/rayfactor.h."""
    import ctypes
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("no C compiler")
    structs = {"rf_solve_opts": _lib.SolveOpts, "rf_timing": _lib.Timing, "rf_host_scene": _lib.HostScene}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "rayfactor.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_solve_opts_defaults():
    import ctypes
    o = _lib.SolveOpts()
    rf.lib().rf_solve_opts_default(ctypes.byref(o))
    assert (o.math, o.rng, o.use_bvs, o.budget, o.n_samples, o.form, o.refine) == (0, 0, 0, 0, 1000000, 1, 1)
    assert (o.rays_per_thread, o.block_size, o.transform, o.refine_margin, o.per_test) == (0, 0, 0, 0.0, 0)


def test_same_tallies_grouping_key():
    """Streaming flushes share a launch only between requests whose options
    give the same tallies (host-only check of the grouping key)."""
    from rayfactorcl_amd.solver import same_tallies
    base = dict(math="fast", rng="seeded", rays_per_emitter=1000, seed=3)
    a = rf.make_opts(**base)
    assert same_tallies(a, rf.make_opts(**base))
    # per-request fields: geometry word, explicit budget, dispatch form
    assert same_tallies(a, rf.make_opts(**{**base, "geometry": 7, "rays_per_emitter": 5000, "form": "per_emitter"}))
    # every option that changes a tally (or the launch shape) separates
    for kw in (dict(refine_margin=0.01), dict(transform="mfma"), dict(refine=False), dict(seed=4),
               dict(math="fp64"), dict(rng="reference"), dict(per_test=1), dict(rays_per_thread=4),
               dict(block_size=128)):
        assert not same_tallies(a, rf.make_opts(**{**base, **kw})), kw
    # the default margin is 2^-14 whether given as 0 or explicitly
    assert same_tallies(a, rf.make_opts(**{**base, "refine_margin": 2.0 ** -14}))
    # margins and transforms do not matter outside the fast form
    c = dict(math="compat", n_samples=1000)
    assert same_tallies(rf.make_opts(**c), rf.make_opts(**c, refine_margin=0.01, transform="mfma"))
    assert not same_tallies(rf.make_opts(**c), rf.make_opts(math="compat", n_samples=2000))
