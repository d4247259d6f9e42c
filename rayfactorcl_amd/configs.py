"""Report of the five workload configurations on one GPU.

For each configuration: throughput (ray-primitive tests/s, rays/s), the
view-factor deviation of the FP32 fast form from an all-FP64 run of the same
rays (same Philox streams), the fraction of individual rays whose nearest hit
differs between the two, and the configuration's own checks (closed forms for
config 1, bounding surface on/off for config 2, streaming latency for
config 5).  Config 4 is bench.py's workload.

    python -m rayfactorcl_amd.configs [--only 1,2,3,5] [--quick]

Prints one JSON object per configuration.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import scenes
from .solver import Context, DeviceScene, Stream, areas, make_opts, solve, trace, vf_report

# (emitter, receiver, closed form, edge allowance, what) for the config-1 scene:
# closed cylinder r=1, H=2 with end discs; coaxial unit discs at spacing 1.
C1_CLOSED = [
    (1, 2, 0.1715729, 0.0, "end disc -> opposite end disc"),
    (1, 0, 0.8284271, 1.5e-3, "end disc -> wall"),
    (0, 1, 0.2071068, 1.5e-3, "wall -> end disc"),
    (0, 0, 0.5857864, 1.5e-3, "wall -> wall"),
    (3, 4, 0.3819660, 0.0, "coaxial unit discs, spacing 1"),
]


def _rate(r) -> dict:
    t = r.timing
    s = max(t["kernel_ms"], 1e-9) * 1e-3
    return {"tests_per_s": t["tests"] / s, "rays_per_s": t["rays"] / s, "kernel_ms": t["kernel_ms"],
            "refined_pairs": t["refined"]}


def _dvf(a, b) -> float:
    """Largest |VF_fast - VF_fp64| over all (emitter, receiver) pairs."""
    rays = np.maximum(a.rays, 1)[:, None]
    return float((np.abs(a.hits - b.hits) / rays).max(initial=0.0))


def _ray_diff(ctx, dev, n, n_wi, bv, seeds=8):
    """Fraction of rays whose nearest primitive differs between fast and FP64,
    over the first n_wi work items of up to `seeds` evenly spaced emitters."""
    diff = total = 0
    for e in np.linspace(0, n - 1, min(n, seeds)).astype(int):
        a = trace(ctx, dev, int(e), 0, n_wi, math="fast", rng="seeded", use_bvs=bv)
        b = trace(ctx, dev, int(e), 0, n_wi, math="fp64", rng="seeded", use_bvs=bv)
        diff += int((a != b).sum())
        total += a.size
    return diff / max(total, 1), total


def _report(vf, area, min_product: float = 1e-3) -> dict:
    """Closure and reciprocity summary of one view-factor matrix."""
    r = vf_report(vf, area, min_product)
    c = r.pop("closure")
    r["closure"] = [round(float(x), 8) for x in c]
    return r


def config1(ctx, quick: bool) -> dict:
    sc = scenes.parse_xml(scenes.config1_xml())
    dev = DeviceScene(ctx, sc.prims, None)
    rays = 1_000_000 if quick else 10_000_000
    out = {"config": 1, "scene": "closed cylinder r=1 H=2 + end discs, coaxial unit discs at spacing 1",
           "primitives": sc.n, "rays_per_emitter": rays}
    res = {}
    area = areas(sc.prims)
    for name, kw in [("compat", dict(math="compat", rng="reference")), ("fast", dict(math="fast", rng="seeded")),
                     ("fp64", dict(math="fp64", rng="seeded"))]:
        r = solve(ctx, dev, rays_per_emitter=rays, form="persistent", **kw)
        res[name] = r
        checks = []
        for (i, j, exact, allow, what) in C1_CLOSED:
            f = float(r.vf[i, j])
            sig = 2.0 * math.sqrt(exact * (1 - exact) / rays)   # 4-ray work items share a block
            checks.append({"pair": [i, j], "what": what, "vf": f, "closed_form": exact, "sigma": sig,
                           "within_4sigma_plus_edge": abs(f - exact) <= 4 * sig + allow})
        out[name] = {**_rate(r), "closed_forms": checks, "report": _report(r.vf, area)}
    out["fast_vs_fp64_max_dvf"] = _dvf(res["fast"], res["fp64"])
    dev.close()
    return out


def config2(ctx, quick: bool) -> dict:
    sc = scenes.parse_xml(scenes.furnace_xml(with_bv=True))
    dev = DeviceScene(ctx, sc.prims, sc.bvs)
    rays = 10_000_000 if quick else 100_000_000
    out = {"config": 2, "scene": "furnace: 180 wall bands + 60 preform bands + neck + annuli + end disc, one bounding surface per group",
           "primitives": sc.n, "bounding_volumes": len(sc.bvs), "rays_per_emitter": rays}
    area = areas(sc.prims)
    for bv in (False, True):
        key = "with_bv" if bv else "without_bv"
        a = solve(ctx, dev, math="fast", rng="seeded", rays_per_emitter=rays, form="persistent", use_bvs=bv)
        b = solve(ctx, dev, math="fp64", rng="seeded", rays_per_emitter=rays, form="persistent", use_bvs=bv)
        frac, nr = _ray_diff(ctx, dev, sc.n, 1 << 16, bv)
        out[key] = {"fast": _rate(a), "fp64": _rate(b), "fast_vs_fp64_max_dvf": _dvf(a, b),
                    "differing_rays": frac, "rays_compared": nr, "fast_report": _report(a.vf, area)}
    dev.close()
    return out


def config3(ctx, quick: bool) -> dict:
    sc = scenes.parse_xml(scenes.soup_xml(1000))
    dev = DeviceScene(ctx, sc.prims, None)
    rays = 100_000 if quick else 1_000_000
    a = solve(ctx, dev, math="fast", rng="seeded", rays_per_emitter=rays, form="persistent")
    b = solve(ctx, dev, math="fp64", rng="seeded", rays_per_emitter=rays, form="persistent")
    c = solve(ctx, dev, math="fast", rng="seeded", rays_per_emitter=rays, form="persistent", refine=False)
    frac, nr = _ray_diff(ctx, dev, sc.n, 1 << 14, False)
    dev.close()
    return {"config": 3, "scene": "random soup of 1000 primitives (cylinder, frustum, disc, annulus, sphere)",
            "primitives": sc.n, "rays_per_emitter": rays, "fast": _rate(a), "fp64": _rate(b),
            "fast_vs_fp64_max_dvf": _dvf(a, b), "fast_unrefined_vs_fp64_max_dvf": _dvf(c, b),
            "differing_rays": frac, "rays_compared": nr}


def _stream_throughput(ctx, reqs, rays: int = 100_000) -> float:
    """Throughput of the streaming service on this request mix: each kind of
    scene (furnace, soup) is timed in one closed batch of all its requests at
    `rays` per emitter, and the mix rate weights the kinds by their share of
    the requests' tests."""
    rate = {}
    for kind in ("furnace", "soup"):
        sel = [(i, sc) for i, (_, sc, _, k) in enumerate(reqs) if k == kind]
        if not sel:
            continue
        st = Stream(ctx, 64 << 30)
        ids = [st.submit(sc.prims, make_opts(math="fast", rng="seeded", rays_per_emitter=rays, form="persistent",
                                             geometry=i)) for i, sc in sel]
        t0 = time.perf_counter()
        st.flush()
        dt = time.perf_counter() - t0
        for rid in ids:
            st.collect(rid)
        st.close()
        rate[kind] = sum(sc.n * sc.n for _, sc in sel) * 4 * ((rays + 3) // 4) / dt
    cost = [(sc.n * sc.n * 4 * ((r + 3) // 4), k) for _, sc, r, k in reqs]
    return sum(c for c, _ in cost) / sum(c / rate[k] for c, k in cost)


def config5(ctx, quick: bool, n_req: int = 0, load: float = 0.7) -> dict:
    """Poisson arrivals at `load` x the measured throughput of the request mix;
    requests of 10-2000 primitives.  Ray budgets are log-U[1e5, 1e7] (1e5-1e8
    in the configuration): the largest request of the full range, 2000
    primitives at 1e8 rays per emitter, is 4e14 tests, minutes of work in one
    flush; 1e7 keeps every request under a few tens of seconds.  The budget
    range is part of the record."""
    n_req = n_req or (30 if quick else 60)
    lo, hi = 1e5, (1e6 if quick else 1e7)
    reqs = scenes.streaming_requests(n_req, rate_hz=1.0, min_rays=lo, max_rays=hi)
    cost = np.array([sc.n * sc.n * (4 * ((r + 3) // 4)) for _, sc, r, _ in reqs], dtype=float)
    throughput = _stream_throughput(ctx, reqs)
    print(f"[config5] calibration: {throughput:.3e} tests/s", file=sys.stderr, flush=True)
    rate = load * throughput / cost.mean()
    arrivals = [t / rate for t, _, _, _ in reqs]
    st = Stream(ctx, 64 << 30)
    t0 = time.perf_counter()
    i = 0
    queued = {}
    lat = []
    rays_done = tests_done = 0
    batches = []
    while i < len(reqs) or queued:
        now = time.perf_counter() - t0
        while i < len(reqs) and arrivals[i] <= now:
            _, sc, r, _ = reqs[i]
            o = make_opts(math="fast", rng="seeded", rays_per_emitter=r, form="persistent", geometry=i)
            queued[st.submit(sc.prims, o)] = (arrivals[i], sc.n)
            i += 1
        if queued:
            n = st.flush()
            done = time.perf_counter() - t0
            batches.append(n)
            for rid, (arr, npr) in list(queued.items()):
                _, rays, _ = st.collect(rid)
                lat.append(done - arr)
                rays_done += int(rays.sum())
                tests_done += int(rays.sum()) * npr
            queued.clear()
            print(f"[config5] {len(lat)}/{n_req} requests done at {done:.1f} s", file=sys.stderr, flush=True)
        elif i < len(reqs):
            time.sleep(max(0.0, min(arrivals[i] - (time.perf_counter() - t0), 0.05)))
    wall = time.perf_counter() - t0
    st.close()
    lat = np.sort(np.asarray(lat))
    return {"config": 5, "requests": n_req, "primitives": "U{10..2000}, furnace or soup",
            "rays_per_emitter": f"log-U[{lo:.0e}, {hi:.0e}]", "offered_load": load,
            "measured_throughput_tests_per_s": throughput,
            "arrival_rate_hz": rate, "wall_s": wall, "rays_per_s": rays_done / wall, "tests_per_s": tests_done / wall,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_max_s": float(lat[-1]), "mean_batch": float(np.mean(batches)), "batches": len(batches)}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--only", default="1,2,3,5")
    ap.add_argument("--quick", action="store_true", help="reduced ray budgets")
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args()
    runs = {1: config1, 2: config2, 3: config3, 5: config5}
    ctx = Context(args.device)
    for k in [int(x) for x in args.only.split(",") if x]:
        t0 = time.perf_counter()
        rec = runs[k](ctx, args.quick)
        rec["seconds"] = time.perf_counter() - t0
        print(json.dumps(rec), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
