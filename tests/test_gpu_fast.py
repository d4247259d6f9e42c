"""The FP32 fast form (hardware rcp/sqrt, FP64 re-solve of ill-conditioned
roots) against the FP64 form on identical Philox streams.  Tolerance: view
factors within 1e-4 of the FP64 result (the precision target), measured both
against the CPU FP64 oracle and, at larger sizes, against the GPU FP64
kernel, which is itself bit-exact with the oracle (test_gpu_parity)."""
import numpy as np
import pytest

import rayfactorcl_amd as rf
from rayfactorcl_amd import scenes

pytestmark = pytest.mark.gpu

VF_TOL = 1e-4


def test_fast_vs_fp64_oracle_config1(ctx, oracle, golden_scene):
    sc = golden_scene("config1")
    dev = rf.DeviceScene(ctx, sc.prims, None)
    n_wi = 1 << 16
    got = rf.solve(ctx, dev, math="fast", rays_per_emitter=4 * n_wi)
    want, rays = oracle.solve(sc.prims, None, n_wi=n_wi, mode="fp64")
    dvf = np.abs(got.hits - want).max() / rays[0]
    print(f"[fast] config1 max |dVF| = {dvf:.2e}", flush=True)
    assert dvf <= VF_TOL
    # ray-for-ray: almost every ray lands on the same primitive
    for e in range(sc.n):
        a = rf.trace(ctx, dev, e, 0, 8192, math="fast")
        b = oracle.trace(sc.prims, None, e, 0, 8192, mode="fp64")
        assert (a != b).mean() < 1e-3


@pytest.mark.parametrize("refine", [True, False])
def test_fast_vs_gpu_fp64_furnace(ctx, refine):
    sc = scenes.parse_xml(scenes.furnace_xml())
    dev = rf.DeviceScene(ctx, sc.prims, None)
    rays = 1 << 18
    a = rf.solve(ctx, dev, math="fast", refine=refine, rays_per_emitter=rays, rng="seeded", form="persistent")
    b = rf.solve(ctx, dev, math="fp64", rays_per_emitter=rays, rng="seeded", form="persistent")
    dvf = np.abs(a.hits - b.hits).max() / rays
    print(f"[fast] furnace refine={refine} max |dVF| = {dvf:.2e}, refined pairs {a.timing['refined']}", flush=True)
    assert dvf <= VF_TOL
    if refine:
        assert a.timing["refined"] > 0


def test_fast_soup_precision(ctx):
    sc = scenes.parse_xml(scenes.soup_xml(300, seed=9))
    dev = rf.DeviceScene(ctx, sc.prims, None)
    rays = 1 << 16
    a = rf.solve(ctx, dev, math="fast", rays_per_emitter=rays, rng="seeded", form="persistent")
    b = rf.solve(ctx, dev, math="fp64", rays_per_emitter=rays, rng="seeded", form="persistent")
    dvf = np.abs(a.hits - b.hits).max() / rays
    print(f"[fast] soup max |dVF| = {dvf:.2e}", flush=True)
    assert dvf <= VF_TOL


@pytest.mark.parametrize("form", ["per_emitter", "persistent", "reference"])
@pytest.mark.parametrize("rpt", [1, 2, 4, 8])
def test_fast_forms_and_shapes_identical(ctx, form, rpt):
    sc = scenes.parse_xml(scenes.furnace_xml(wall_bands=40, preform_bands=12))
    dev = rf.DeviceScene(ctx, sc.prims, None)
    base = rf.solve(ctx, dev, math="fast", rays_per_emitter=20000, rng="seeded", form="persistent",
                    rays_per_thread=4)
    got = rf.solve(ctx, dev, math="fast", rays_per_emitter=20000, rng="seeded", form=form, rays_per_thread=rpt)
    assert np.array_equal(got.hits, base.hits)


def _ring_zoo_xml() -> str:
    """Discs and annuli inside a cylinder wall (r 3, z 0-4): four whose plane is
    z = const and whose x-y is a similarity of world x-y (off-centre, rotated
    about z, flipped by 180 degrees about x: the fast scan tests these in world
    coordinates), one tilted disc and one elliptical disc (the object-frame
    test)."""
    P = scenes._prim
    body = "".join([
        P("cylinderSurface", bounding=True, small=1.0, scale=(3.0, 3.0, 4.0)),
        P("disc", scale=(0.8, 0.8, 1.0), translate=(0.5, -0.3, 1.0)),
        P("annulus", small=0.3, large=1.2, rotate=(37.0, (0.0, 0.0, 1.0)), translate=(-0.7, 0.4, 2.5)),
        P("disc", scale=(1.5, 1.5, 1.0), rotate=scenes.FLIP_X, translate=(0.0, 0.0, 3.5)),
        P("annulus", small=0.5, large=2.0, rotate=scenes.FLIP_X, translate=(0.2, 0.1, 0.2)),
        P("disc", scale=(0.6, 0.6, 1.0), rotate=(20.0, (1.0, 0.0, 0.0)), translate=(1.0, 1.0, 2.0)),
        P("disc", scale=(1.0, 0.5, 1.0), translate=(-1.0, -1.0, 1.5)),
    ])
    return scenes._doc("discs and annuli, world z-planes and others", body)


def test_world_plane_rings_ray_for_ray(ctx, oracle):
    """The fast scan's world-coordinate test of z-plane discs and annuli
    against the FP64 oracle's object-frame test (RayFactorGPU.cl:126-187),
    ray for ray, from every emitter; and the flops it reports: 11 per ray per
    world z-plane ring plus one reciprocal per ray, against the reference
    formula's 42 in the MFMA form, which tests every ring in its object
    frame."""
    sc = scenes.parse_xml(_ring_zoo_xml())
    dev = rf.DeviceScene(ctx, sc.prims, None)
    diff = total = zr = 0
    for e in range(sc.n):
        a = rf.trace(ctx, dev, e, 0, 8192, math="fast")
        b = oracle.trace(sc.prims, None, e, 0, 8192, mode="fp64")
        assert (a != b).mean() < 1e-3, e
        zr += int(np.isin(b, [1, 2, 3, 4]).sum())
        diff += int((a != b).sum())
        total += a.size
    print(f"[rings] {diff} of {total} rays differ from the FP64 oracle; {zr} hit a z-plane ring", flush=True)
    assert zr > 0.05 * total                # the z-plane rings are in the rays' way
    assert diff <= 1e-4 * total
    kw = dict(math="fast", rng="seeded", rays_per_emitter=1 << 14, form="persistent", per_test=1)
    v = rf.solve(ctx, dev, transform="valu", **kw).timing
    m = rf.solve(ctx, dev, transform="mfma", **kw).timing
    assert v["rays"] == m["rays"]
    assert m["flops"] - v["flops"] == v["rays"] * (4 * (42 - 11) - 1)


def test_fast_bv_matches_fp64_bv(ctx):
    sc = scenes.parse_xml(scenes.furnace_xml(with_bv=True))
    dev = rf.DeviceScene(ctx, sc.prims, sc.bvs)
    rays = 1 << 16
    a = rf.solve(ctx, dev, math="fast", use_bvs=True, rays_per_emitter=rays, rng="seeded", form="persistent")
    b = rf.solve(ctx, dev, math="fp64", use_bvs=True, rays_per_emitter=rays, rng="seeded", form="persistent")
    assert np.abs(a.hits - b.hits).max() / rays <= VF_TOL
    # executed tests are counted on the device and are fewer than brute force
    assert 0 < a.timing["tests"] < rays * sc.n * sc.n


def test_fp64_bv_matches_oracle_furnace(ctx, oracle):
    sc = scenes.parse_xml(scenes.furnace_xml(with_bv=True))
    dev = rf.DeviceScene(ctx, sc.prims, sc.bvs)
    n_wi = 64
    a = rf.solve(ctx, dev, math="fp64", use_bvs=True, rays_per_emitter=4 * n_wi)
    b, _ = oracle.solve(sc.prims, sc.bvs, n_wi=n_wi, mode="fp64", use_bvs=True)
    assert np.array_equal(a.hits, b)
    c = rf.solve(ctx, dev, math="compat", use_bvs=True, rays_per_emitter=4 * n_wi)
    d, _ = oracle.solve(sc.prims, sc.bvs, n_wi=n_wi, mode="compat", use_bvs=True)
    assert np.array_equal(c.hits, d)


# ---------------------------------------------------------------------------
# Ray transform on the matrix cores (v_mfma_f32_16x16x4f32)

@pytest.mark.parametrize("refine", [True, False])
def test_mfma_vs_gpu_fp64_furnace(ctx, refine):
    sc = scenes.parse_xml(scenes.furnace_xml())
    dev = rf.DeviceScene(ctx, sc.prims, None)
    rays = 1 << 18
    a = rf.solve(ctx, dev, math="fast", refine=refine, rays_per_emitter=rays, rng="seeded", form="persistent",
                 transform="mfma")
    b = rf.solve(ctx, dev, math="fp64", rays_per_emitter=rays, rng="seeded", form="persistent")
    dvf = np.abs(a.hits - b.hits).max() / rays
    print(f"[mfma] furnace refine={refine} max |dVF| = {dvf:.2e}, refined pairs {a.timing['refined']}", flush=True)
    assert dvf <= VF_TOL
    assert np.array_equal(a.rays, b.rays)
    if refine:
        assert a.timing["refined"] > 0


def test_mfma_soup_and_config1_precision(ctx, golden_scene):
    for sc in (scenes.parse_xml(scenes.soup_xml(300, seed=9)), golden_scene("config1")):
        dev = rf.DeviceScene(ctx, sc.prims, None)
        rays = 1 << 16
        a = rf.solve(ctx, dev, math="fast", rays_per_emitter=rays, rng="seeded", form="persistent",
                     transform="mfma")
        b = rf.solve(ctx, dev, math="fp64", rays_per_emitter=rays, rng="seeded", form="persistent")
        dvf = np.abs(a.hits - b.hits).max() / rays
        print(f"[mfma] {sc.n} primitives max |dVF| = {dvf:.2e}", flush=True)
        assert dvf <= VF_TOL


@pytest.mark.parametrize("block", [64, 128, 256])
def test_mfma_forms_identical(ctx, block):
    """Persistent queue and per-emitter launches of the MFMA kernel hand the
    same rays to the same arithmetic: identical tallies for any group size."""
    sc = scenes.parse_xml(scenes.furnace_xml(wall_bands=40, preform_bands=12))
    dev = rf.DeviceScene(ctx, sc.prims, None)
    kw = dict(math="fast", rays_per_emitter=20000, rng="seeded", transform="mfma")
    base = rf.solve(ctx, dev, form="persistent", block_size=256, **kw)
    got = rf.solve(ctx, dev, form="per_emitter", block_size=block, **kw)
    assert np.array_equal(got.hits, base.hits)


def test_mfma_trace_matches_tally(ctx):
    sc = scenes.parse_xml(scenes.furnace_xml(wall_bands=20, preform_bands=6))
    dev = rf.DeviceScene(ctx, sc.prims, None)
    n_wi = 3000                      # not a multiple of the 16-ray MFMA set: ragged last pass
    res = rf.solve(ctx, dev, math="fast", rays_per_emitter=4 * n_wi, rng="seeded", form="per_emitter",
                   transform="mfma")
    for e in (0, sc.n // 2, sc.n - 1):
        idx = rf.trace(ctx, dev, e, 0, n_wi, math="fast", rng="seeded", transform="mfma")
        assert idx.shape == (4 * n_wi,)
        assert ((idx >= -1) & (idx < sc.n)).all()
        assert np.array_equal(np.bincount(idx[idx >= 0], minlength=sc.n), res.hits[e])


# ---------------------------------------------------------------------------
# Re-solve lists that fill up: a wide flag margin flags most rim crossings, so
# lanes drop pairs and re-scan; results must not depend on the launch shape.

WIDE = 0.25


@pytest.mark.parametrize("form", ["per_emitter", "persistent", "reference"])
@pytest.mark.parametrize("rpt", [1, 2, 4, 8])
def test_wide_margin_forms_and_shapes_identical(ctx, form, rpt):
    sc = scenes.parse_xml(scenes.furnace_xml(wall_bands=40, preform_bands=12))
    dev = rf.DeviceScene(ctx, sc.prims, None)
    kw = dict(math="fast", rays_per_emitter=20000, rng="seeded", refine_margin=WIDE, transform="valu")
    base = rf.solve(ctx, dev, form="persistent", rays_per_thread=4, **kw)
    got = rf.solve(ctx, dev, form=form, rays_per_thread=rpt, **kw)
    assert np.array_equal(got.hits, base.hits)
    # every flagged pair is re-solved exactly once, whatever the shape
    assert got.timing["refined"] == base.timing["refined"]
    if rpt == 4 and form == "persistent":
        # lists overflowed and the re-scan path ran
        assert base.timing["rescans"] > 0
        print(f"[wide] refined {base.timing['refined']}, re-scans {base.timing['rescans']}", flush=True)


@pytest.mark.parametrize("transform", ["valu", "mfma"])
def test_wide_margin_precision(ctx, transform):
    """Re-solving far more pairs than needed only moves the result closer to
    FP64; the MFMA kernel's re-scan is checked against its per-emitter form."""
    sc = scenes.parse_xml(scenes.furnace_xml())
    dev = rf.DeviceScene(ctx, sc.prims, None)
    rays = 1 << 16
    kw = dict(math="fast", rays_per_emitter=rays, rng="seeded", transform=transform)
    a = rf.solve(ctx, dev, form="persistent", refine_margin=WIDE, **kw)
    d = rf.solve(ctx, dev, form="persistent", **kw)
    b = rf.solve(ctx, dev, math="fp64", rays_per_emitter=rays, rng="seeded", form="persistent")
    dvf = np.abs(a.hits - b.hits).max() / rays
    print(f"[wide] {transform} max |dVF| = {dvf:.2e} (default margin "
          f"{np.abs(d.hits - b.hits).max() / rays:.2e}), refined {a.timing['refined']} vs {d.timing['refined']}, "
          f"re-scans {a.timing['rescans']}", flush=True)
    assert dvf <= VF_TOL
    assert a.timing["refined"] > 10 * d.timing["refined"]
    assert a.timing["rescans"] > 0
    c = rf.solve(ctx, dev, form="per_emitter", refine_margin=WIDE, **kw)
    assert np.array_equal(a.hits, c.hits)


# ---------------------------------------------------------------------------
# Bounding surfaces: the binned kernel (rays bucketed by nearest surface) in the
# queue forms, the one-ray-per-thread kernel in the reference form and for
# surfaces that ask for the intermediate check.

def _furnace_bv():
    return scenes.parse_xml(scenes.furnace_xml(with_bv=True))


def test_bv_binned_matches_reference_form_and_fp64(ctx):
    sc = _furnace_bv()
    dev = rf.DeviceScene(ctx, sc.prims, sc.bvs)
    rays = 1 << 16
    kw = dict(math="fast", use_bvs=True, rays_per_emitter=rays, rng="seeded")
    a = rf.solve(ctx, dev, form="persistent", **kw)
    r = rf.solve(ctx, dev, form="reference", **kw)
    d = rf.solve(ctx, dev, math="fp64", use_bvs=True, rays_per_emitter=rays, rng="seeded", form="persistent")
    assert np.abs(a.hits - d.hits).max() / rays <= VF_TOL
    # identical arithmetic per pair; only exact FP32 ties across primitive types may resolve differently
    moved = int(np.abs(a.hits - r.hits).sum())
    print(f"[bvbin] rays resolved differently from the reference form: {moved // 2} of {rays * sc.n}", flush=True)
    assert moved <= 20
    # the reference form tests every child of the winning surface; the binned
    # form searches the children's band stacks
    assert a.timing["tests_brute_force"] == r.timing["tests_brute_force"] == rays * sc.n * sc.n
    assert 0 < a.timing["tests"] < r.timing["tests"] < rays * sc.n * sc.n


@pytest.mark.parametrize("refine", [True, False])
def test_bv_binned_forms_identical_and_traces(ctx, refine):
    sc = _furnace_bv()
    dev = rf.DeviceScene(ctx, sc.prims, sc.bvs)
    n_wi = 5000                        # ragged last pass (1024 rays per pass)
    kw = dict(math="fast", use_bvs=True, rays_per_emitter=4 * n_wi, rng="seeded", refine=refine)
    a = rf.solve(ctx, dev, form="persistent", **kw)
    b = rf.solve(ctx, dev, form="per_emitter", **kw)
    assert np.array_equal(a.hits, b.hits)
    for e in (0, 100, sc.n - 1):
        tr = rf.trace(ctx, dev, e, 0, n_wi, math="fast", use_bvs=True, rng="seeded", refine=refine)
        assert np.array_equal(np.bincount(tr[tr >= 0], minlength=sc.n), a.hits[e])


def test_bv_intermediate_check_uses_per_ray_kernel(ctx, oracle):
    """Surfaces with the intermediate check (settings == 1) keep the per-ray
    kernel; it agrees with the FP64 form, and compat agrees with the oracle."""
    sc = _furnace_bv()
    bvs = sc.bvs.copy()
    bvs["settings"][:2] = 1
    dev = rf.DeviceScene(ctx, sc.prims, bvs)
    rays = 1 << 14
    a = rf.solve(ctx, dev, math="fast", use_bvs=True, rays_per_emitter=rays, rng="seeded", form="persistent")
    d = rf.solve(ctx, dev, math="fp64", use_bvs=True, rays_per_emitter=rays, rng="seeded", form="persistent")
    assert np.abs(a.hits - d.hits).max() / rays <= VF_TOL
    c = rf.solve(ctx, dev, math="compat", use_bvs=True, rays_per_emitter=256)
    want, _ = oracle.solve(sc.prims, bvs, n_wi=64, mode="compat", use_bvs=True)
    assert np.array_equal(c.hits, want)


def test_bv_binned_empty_and_overlapping_children(ctx, oracle):
    """A surface with no children swallows its rays (reference semantics), and
    child ranges may overlap; the binned kernel agrees with the per-ray kernel
    of the reference form, and compat with the oracle."""
    sc = _furnace_bv()
    bvs = sc.bvs.copy()
    bvs["objEndIndex"][2] = bvs["objStartIndex"][2]            # no children
    bvs["objStartIndex"][1] = max(0, int(bvs["objStartIndex"][1]) - 5)   # overlaps its neighbour's range
    dev = rf.DeviceScene(ctx, sc.prims, bvs)
    rays = 1 << 14
    kw = dict(math="fast", use_bvs=True, rays_per_emitter=rays, rng="seeded")
    a = rf.solve(ctx, dev, form="persistent", **kw)
    r = rf.solve(ctx, dev, form="reference", **kw)
    assert int(np.abs(a.hits - r.hits).sum()) <= 10
    c = rf.solve(ctx, dev, math="compat", use_bvs=True, rays_per_emitter=256)
    want, _ = oracle.solve(sc.prims, bvs, n_wi=64, mode="compat", use_bvs=True)
    assert np.array_equal(c.hits, want)


def _cylinder_zoo_xml() -> str:
    """Cylinder emitters inside a closed wall: uniform cross-sections (the
    emitter frame built once on the host: upright, rotated about z, tilted),
    an elliptical one and a sheared one (the frame built per ray), and a
    frustum."""
    P = scenes._prim
    body = "".join([
        P("cylinderSurface", bounding=True, small=1.0, scale=(3.0, 3.0, 4.0)),
        P("disc", scale=(3.0, 3.0, 1.0)),
        P("disc", scale=(3.0, 3.0, 1.0), rotate=scenes.FLIP_X, translate=(0.0, 0.0, 4.0)),
        P("cylinderSurface", small=1.0, scale=(0.4, 0.4, 1.0), translate=(1.0, 0.5, 0.5)),
        P("cylinderSurface", small=1.0, scale=(0.3, 0.3, 0.8), rotate=(37.0, (0.0, 0.0, 1.0)),
          translate=(-1.2, 0.3, 1.5)),
        P("cylinderSurface", small=1.0, scale=(0.3, 0.3, 1.0), rotate=(30.0, (1.0, 0.0, 0.0)),
          translate=(0.5, -1.2, 1.2)),
        P("cylinderSurface", small=1.0, scale=(0.5, 0.25, 1.0), translate=(-0.8, -1.0, 2.2)),
        P("cylinderSurface", small=1.0, scale=(0.3, 0.3, 1.0), rotate=(25.0, (1.0, 1.0, 0.0)),
          translate=(1.5, 1.5, 2.5)),
        P("frustum", small=0.5, scale=(0.5, 0.5, 0.8), translate=(0.0, 1.6, 0.3)),
    ])
    return scenes._doc("cylinder emitters: host-built and per-ray frames", body)


def test_cylinder_emitter_frames_ray_for_ray(ctx, oracle):
    """Cylinder emitters whose frame is built once on the host (uniform
    cross-section, any rotation) and per ray (elliptical, the rest) against
    the FP64 oracle's per-ray Gram-Schmidt frame (RayFactorGPU.cl:347-365 for
    the sampling), ray for ray: the same rays up to FP32 rounding, so almost
    every ray lands on the same primitive from every emitter."""
    sc = scenes.parse_xml(_cylinder_zoo_xml())
    dev = rf.DeviceScene(ctx, sc.prims, None)
    diff = total = 0
    for e in range(sc.n):
        a = rf.trace(ctx, dev, e, 0, 8192, math="fast")
        b = oracle.trace(sc.prims, None, e, 0, 8192, mode="fp64")
        assert (a != b).mean() < 1e-3, e
        assert (b >= 0).mean() > 0.99, e          # the enclosure is closed
        diff += int((a != b).sum())
        total += a.size
    print(f"[frames] {diff} of {total} rays differ from the FP64 oracle", flush=True)
    assert diff <= 2e-4 * total
    dev.close()
