"""Writes the golden fixtures of tests/golden/ with the programs built from the reference sources.

Runs ONLY in a container that holds the reference sources (``make -C oracle ref``
builds ``oracle/_ref/`` from them).  Outputs, all small and committed:

* ``scenes/<name>.rfscene`` -- the Primitive / BoundingVolume bytes that the
  reference's own loader (RayFactorCL/Scene.cpp, compiled from its sources)
  produces for every XML under ``xml/`` (the reference's sample-geometries
  plus ``config1.xml``).  Fields the reference leaves uninitialised read as
  zero (the dump driver zero-fills every allocation).
* ``philox_vectors.json`` -- Philox4x32-10 outputs of the reference's
  Random123 header for the published known-answer inputs, the reference
  kernel's per-work-item stream (key {gid, 0xdecafbad}) and a set of
  seeded-mode blocks.
"""
from __future__ import annotations

import json
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
REF = ROOT / "oracle" / "_ref"


def main() -> None:
    subprocess.run(["make", "-C", str(ROOT / "oracle"), "ref"], check=True, capture_output=True)
    out = HERE / "scenes"
    out.mkdir(exist_ok=True)
    for xml in sorted((HERE / "xml").glob("*.xml")):
        dst = out / (xml.stem + ".rfscene")
        subprocess.run([str(REF / "ref_scene_dump"), str(xml), str(dst)], check=True)

    vectors = []
    # published Random123 known-answer inputs (ctr[4], key[2])
    kat_inputs = [
        ([0, 0, 0, 0], [0, 0]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]),
    ]
    for c, k in kat_inputs:
        vectors.append({"kind": "kat", "ctr": c, "key": k})
    # reference kernel stream: ctr {1, f00dcafe, deadbeef, beeff00d}, key {gid, decafbad}
    for gid in list(range(64)) + [999935, 2**20 + 7, 2**31 - 1, 2**32 - 1]:
        vectors.append({"kind": "reference", "gid": gid,
                        "ctr": [1, 0xF00DCAFE, 0xDEADBEEF, 0xBEEFF00D], "key": [gid, 0xDECAFBAD]})
    # seeded stream: key {wi_lo, seed}, ctr {wi_hi, emitter, geometry, beeff00d}
    for wi, seed, em, geo in [(0, 0x52465331, 0, 0), (12345, 0x52465331, 7, 3), (2**33 + 5, 99, 244, 255),
                              (2**32 - 1, 1, 1, 1)]:
        vectors.append({"kind": "seeded", "wi": wi, "seed": seed, "emitter": em, "geometry": geo,
                        "ctr": [wi >> 32, em, geo, 0xBEEFF00D], "key": [wi & 0xFFFFFFFF, seed]})
    lines = "".join("%x %x %x %x %x %x\n" % (*v["ctr"], *v["key"]) for v in vectors)
    res = subprocess.run([str(REF / "ref_philox_dump")], input=lines.encode(), check=True, capture_output=True)
    outs = res.stdout.decode().split("\n")
    for v, o in zip(vectors, outs):
        v["out"] = [int(x, 16) for x in o.split()]
    (HERE / "philox_vectors.json").write_text(json.dumps(vectors, indent=1) + "\n")


if __name__ == "__main__":
    main()
