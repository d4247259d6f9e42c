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
from ._lib import RF_ERR_CAPACITY, RayFactorError
from .solver import Context, DeviceScene, Stream, areas, make_opts, solve, trace, vf_report

EPS = 1e-3             # self-intersection cut-off (RayFactorCL/RayFactorGPU.cl:7)


def seam_loss(edge_len: float, area_i: float, alpha: float) -> float:
    """First-order view-factor loss of emitter i toward a surface that meets it
    at dihedral angle alpha along an edge of length edge_len: rays that reach
    the neighbour sooner than EPS pass through it.  A point at distance x from
    the edge is x sin(alpha) from the neighbour's plane, so averaged over x and
    over cosine-weighted directions d the loss is
    L EPS E[max(0, d.v)] / (A_i sin(alpha)), v at angle alpha from the normal,
    with E[max(0, d.v)] = 2((pi - alpha) cos(alpha) + sin(alpha)) / (3 pi)."""
    k = 2.0 * ((math.pi - alpha) * math.cos(alpha) + math.sin(alpha)) / (3.0 * math.pi)
    return edge_len * EPS * k / (area_i * math.sin(alpha))


# (emitter, receiver, closed form, seam loss, what) for the config-1 scene:
# closed cylinder r=1, H=2 (wall area 4 pi) with end discs (pi) meeting it at
# right angles along rims of 2 pi; coaxial unit discs at spacing 1.
C1_CLOSED = [
    (1, 2, 0.1715729, 0.0, "end disc -> opposite end disc"),
    (1, 0, 0.8284271, seam_loss(2 * math.pi, math.pi, math.pi / 2), "end disc -> wall"),
    (0, 1, 0.2071068, seam_loss(2 * math.pi, 4 * math.pi, math.pi / 2), "wall -> end disc"),
    (0, 0, 0.5857864, 0.0, "wall -> wall"),
    (3, 4, 0.3819660, 0.0, "coaxial unit discs, spacing 1"),
]
REMAINDER = 1e-5       # O(EPS^2) remainder of the seam-loss expansion for unit-size scenes


def _vf_se(idx: np.ndarray, j: int) -> tuple[float, float]:
    """View factor toward j and its standard error from a per-ray trace (the 4
    rays of a work item share one random block: variance over work items)."""
    k = (idx.reshape(-1, 4) == j).sum(axis=1).astype(np.float64)
    return float(k.mean() / 4.0), float(k.std(ddof=1) / 4.0 / math.sqrt(len(k)))


def cpu_baseline(prims, bvs, emitters, seconds: float, threads: int, use_bvs: bool = False) -> dict:
    """The CPU oracle (compat FP32 restatement of the reference kernel, brute
    force over every primitive) on the host cores: the given emitters, ray
    budget calibrated so the timed run takes about `seconds`."""
    import os
    import sys as _sys
    from pathlib import Path
    _sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
    import oracle as O
    n = len(prims)
    e0, e1 = int(min(emitters)), int(max(emitters)) + 1
    n_wi = 64
    for budget in (0.2, 1.0):
        t0 = time.perf_counter()
        O.solve(prims, bvs, n_wi=n_wi, mode="compat", rng="seeded", threads=threads, emitters=(e0, e1),
                use_bvs=use_bvs)
        dt = max(time.perf_counter() - t0, 1e-6)
        n_wi = max(64, int(n_wi * budget / dt))
    n_wi = max(64, int(n_wi * seconds))
    t0 = time.perf_counter()
    _, rays = O.solve(prims, bvs, n_wi=n_wi, mode="compat", rng="seeded", threads=threads, emitters=(e0, e1),
                      use_bvs=use_bvs)
    dt = time.perf_counter() - t0
    total = int(rays.sum())
    return {"rays_per_s": total / dt, "tests_per_s": total * n / dt if not use_bvs else None,
            "threads": threads, "seconds": dt,
            "sample": f"emitters {e0}..{e1 - 1}, {4 * n_wi} rays each, {os.cpu_count()} host CPUs"}


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


def config1(ctx, quick: bool, cpu_seconds: float = 10.0, threads: int = 16) -> dict:
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
        tr = {e: trace(ctx, dev, e, 0, rays // 4, **kw) for e in (0, 1, 3)}
        checks = []
        for (i, j, exact, loss, what) in C1_CLOSED:
            f, se = _vf_se(tr[i], j)
            checks.append({"pair": [i, j], "what": what, "vf": f, "closed_form": exact, "seam_loss": loss,
                           "se": se, "z": (f - (exact - loss)) / se,
                           "within_3se": abs(f - (exact - loss)) <= 3 * se + REMAINDER})
        out[name] = {**_rate(r), "closed_forms": checks, "report": _report(r.vf, area)}
    out["fast_vs_fp64_max_dvf"] = _dvf(res["fast"], res["fp64"])
    if cpu_seconds > 0:
        out["cpu_baseline"] = cpu_baseline(sc.prims, None, range(sc.n), cpu_seconds, threads)
        out["cpu_baseline"]["gpu_compat_over_cpu"] = out["compat"]["rays_per_s"] / out["cpu_baseline"]["rays_per_s"]
    dev.close()
    return out


def config2(ctx, quick: bool, cpu_seconds: float = 10.0, threads: int = 16) -> dict:
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
        # the reference arithmetic on the GPU, band stacks searched and every test in full
        rc = 1 << 22
        c = solve(ctx, dev, math="compat", rng="seeded", rays_per_emitter=rc, form="persistent", use_bvs=bv)
        d = solve(ctx, dev, math="compat", rng="seeded", rays_per_emitter=rc, form="persistent", use_bvs=bv,
                  per_test=1)
        frac, nr = _ray_diff(ctx, dev, sc.n, 1 << 16, bv)
        out[key] = {"fast": _rate(a), "fp64": _rate(b), "compat": _rate(c), "compat_per_test": _rate(d),
                    "fast_vs_fp64_max_dvf": _dvf(a, b),
                    "differing_rays": frac, "rays_compared": nr, "fast_report": _report(a.vf, area)}
        if cpu_seconds > 0:
            em = np.linspace(0, sc.n - 1, 8).astype(int)
            cb = cpu_baseline(sc.prims, sc.bvs if bv else None, range(em[0], em[0] + 8), cpu_seconds, threads,
                              use_bvs=bv)
            cb["gpu_compat_per_test_over_cpu"] = out[key]["compat_per_test"]["rays_per_s"] / cb["rays_per_s"]
            cb["gpu_compat_over_cpu"] = out[key]["compat"]["rays_per_s"] / cb["rays_per_s"]
            out[key]["cpu_baseline"] = cb
    dev.close()
    return out


def config3(ctx, quick: bool, cpu_seconds: float = 10.0, threads: int = 16) -> dict:
    sc = scenes.parse_xml(scenes.soup_xml(1000))
    dev = DeviceScene(ctx, sc.prims, None)
    rays = 100_000 if quick else 1_000_000
    a = solve(ctx, dev, math="fast", rng="seeded", rays_per_emitter=rays, form="persistent")
    b = solve(ctx, dev, math="fp64", rng="seeded", rays_per_emitter=rays, form="persistent")
    c = solve(ctx, dev, math="fast", rng="seeded", rays_per_emitter=rays, form="persistent", refine=False)
    d = solve(ctx, dev, math="compat", rng="seeded", rays_per_emitter=rays, form="persistent")
    # the fast form without its bounding-box hierarchy: every primitive tested
    e = solve(ctx, dev, math="fast", rng="seeded", rays_per_emitter=rays, form="persistent", hierarchy=False)
    frac, nr = _ray_diff(ctx, dev, sc.n, 1 << 14, False)
    out = {"config": 3, "scene": "random soup of 1000 primitives (cylinder, frustum, disc, annulus, sphere)",
           "primitives": sc.n, "rays_per_emitter": rays, "fast": _rate(a), "fp64": _rate(b), "compat": _rate(d),
           "fast_every_primitive": _rate(e), "fast_hierarchy_equals_every_primitive": bool(np.array_equal(a.hits, e.hits)),
           "fast_vs_fp64_max_dvf": _dvf(a, b), "fast_unrefined_vs_fp64_max_dvf": _dvf(c, b),
           "differing_rays": frac, "rays_compared": nr}
    if cpu_seconds > 0:
        # soup cylinders share no rows: GPU and CPU both test every primitive
        out["cpu_baseline"] = cpu_baseline(sc.prims, None, range(32), cpu_seconds, threads)
        out["cpu_baseline"]["gpu_compat_over_cpu"] = out["compat"]["rays_per_s"] / out["cpu_baseline"]["rays_per_s"]
    dev.close()
    return out


def _stream_calibrate(ctx, reqs, rays: int = 100_000):
    """Throughput of the streaming service on this request mix, in the library's
    estimated ray-primitive evaluations per second, and each request's
    estimated cost at its own budget.  Each kind of scene (furnace, soup) is
    timed in one closed batch of all its requests at `rays` per emitter; a
    request's cost is linear in its ray budget, so its cost at `rays`
    (read from the queue) scales to its own.  The mix rate weights the kinds by
    their share of the requests' cost."""
    rate, cost = {}, [0.0] * len(reqs)
    cal_rays = 4 * ((rays + 3) // 4)
    for kind in ("furnace", "soup"):
        sel = [(i, sc) for i, (_, sc, _, k) in enumerate(reqs) if k == kind]
        if not sel:
            continue
        st = Stream(ctx, 64 << 30)
        ids = []
        for i, sc in sel:
            q0 = st.queued_cost()
            ids.append(st.submit(sc.prims, make_opts(math="fast", rng="seeded", rays_per_emitter=rays,
                                                     form="persistent", geometry=i)))
            r_i = reqs[i][2]
            cost[i] = (st.queued_cost() - q0) * (4 * ((r_i + 3) // 4)) / cal_rays
        total = st.queued_cost()
        t0 = time.perf_counter()
        st.flush()
        dt = time.perf_counter() - t0
        for rid in ids:
            st.collect(rid)
        st.close()
        rate[kind] = total / dt
    mix = sum(cost) / sum(c / rate[k] for c, (_, _, _, k) in zip(cost, reqs))
    return mix, np.asarray(cost)


def config5(ctx, quick: bool, n_req: int = 0, load: float = 0.7, queue_s: float = 60.0,
            launch_s: float = 2.0) -> dict:
    """Poisson arrivals at `load` x the measured throughput of the request mix;
    requests of 10-2000 primitives with ray budgets log-U[1e5, 1e8] per emitter
    (quick: 1e5-1e6).  Admission: HBM footprint, and a queued-cost limit of
    `queue_s` seconds of work at the measured throughput; a request turned away
    waits in the client's backlog (first come, first served) and is offered
    again after each flush, so its latency counts from its arrival.  Flushes
    pack requests cheapest first into launches of about `launch_s` seconds of
    work, one launch per step; a request costlier than that runs in chunks of
    emitters, and the loop admits arrivals between launches.  Cost is the
    library's estimate of executed ray-primitive evaluations."""
    n_req = n_req or (24 if quick else 40)
    lo, hi = 1e5, (1e6 if quick else 1e8)
    reqs = scenes.streaming_requests(n_req, rate_hz=1.0, min_rays=lo, max_rays=hi)
    throughput, est = _stream_calibrate(ctx, reqs)
    print(f"[config5] calibration: {throughput:.3e} evaluations/s, largest request {est.max():.2e}",
          file=sys.stderr, flush=True)
    st = Stream(ctx, 64 << 30)
    st.set_cost_limits(max_queued_tests=queue_s * throughput, max_launch_tests=launch_s * throughput)
    rate = load * throughput / est.mean()
    arrivals = [t / rate for t, _, _, _ in reqs]
    t0 = time.perf_counter()
    i = 0
    backlog = []                                 # arrived, not yet admitted (request index)
    queued = {}
    lat = []
    rays_done = tests_done = launches = 0
    held = set()
    batches = []
    while i < len(reqs) or backlog or queued:
        now = time.perf_counter() - t0
        while i < len(reqs) and arrivals[i] <= now:
            backlog.append(i)
            i += 1
        while backlog:
            k = backlog[0]
            _, sc, r, _ = reqs[k]
            o = make_opts(math="fast", rng="seeded", rays_per_emitter=r, form="persistent", geometry=k)
            try:
                queued[st.submit(sc.prims, o)] = (arrivals[k], sc.n, time.perf_counter() - t0)
                backlog.pop(0)
            except RayFactorError as e:          # admission: over the cost (or footprint) limit
                if e.code != RF_ERR_CAPACITY:
                    raise
                held.add(k)
                break
        if queued:
            # one launch, then back to the arrivals: a request that arrives
            # while a large one runs in chunks can go before its next chunk
            n = st.step()
            launches += 1
            if n:
                batches.append(n)
            for rid, (arr, npr, sub) in list(queued.items()):
                try:
                    _, rays, lat_ms = st.collect(rid)
                except RayFactorError:
                    continue                             # not completed yet
                del queued[rid]
                lat.append(sub - arr + lat_ms * 1e-3)      # backlog wait + the library's submit -> done
                rays_done += int(rays.sum())
                tests_done += int(rays.sum()) * npr
            print(f"[config5] launch {launches}: {len(lat)}/{n_req} requests done at "
                  f"{time.perf_counter() - t0:.1f} s (queued {len(queued)}, backlog {len(backlog)})",
                  file=sys.stderr, flush=True)
        elif i < len(reqs):
            time.sleep(max(0.0, min(arrivals[i] - (time.perf_counter() - t0), 0.05)))
    wall = time.perf_counter() - t0
    st.close()
    lat = np.sort(np.asarray(lat))
    return {"config": 5, "requests": n_req, "primitives": "U{10..2000}, furnace or soup",
            "rays_per_emitter": f"log-U[{lo:.0e}, {hi:.0e}]", "offered_load": load,
            "measured_throughput_evaluations_per_s": throughput, "cost_limit_s": queue_s, "launch_limit_s": launch_s,
            "arrival_rate_hz": rate, "wall_s": wall, "rays_per_s": rays_done / wall,
            "brute_force_equivalent_tests_per_s": tests_done / wall,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_max_s": float(lat[-1]), "launches": launches,
            "mean_requests_completed_per_completing_launch": float(np.mean(batches)),
            "requests_held_back_by_cost_admission": len(held),
            "largest_request_evaluations": float(est.max()), "mean_request_evaluations": float(est.mean())}


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
