"""Scene input: the reference XML format (through the native loader that
replaces RayFactorCL/Scene.cpp) and synthetic generators of the benchmark
workloads.

Generators emit XML in the reference format and parse it with the same
loader, so every synthetic scene goes through the reference's affine and
rayDensity rules (Scene.cpp:296-390).
"""
from __future__ import annotations

import math
from ctypes import byref, c_char
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from ._lib import BV_DTYPE, PRIM_DTYPE, HostScene, RayFactorError, lib


@dataclass
class SceneData:
    prims: np.ndarray          # PRIM_DTYPE records (112 B each)
    bvs: np.ndarray            # BV_DTYPE records (72 B each)
    globalRayDensity: float
    useBoundingVolumes: bool

    @property
    def n(self) -> int:
        return len(self.prims)


def _from_host(hs: HostScene) -> SceneData:
    n, nb = hs.n_prims, hs.n_bvs
    prims = np.zeros(n, dtype=PRIM_DTYPE)
    bvs = np.zeros(nb, dtype=BV_DTYPE)
    if n:
        prims[:] = np.frombuffer((c_char * (n * 112)).from_address(hs.prims), dtype=PRIM_DTYPE)
    if nb:
        bvs[:] = np.frombuffer((c_char * (nb * 72)).from_address(hs.bvs), dtype=BV_DTYPE)
    return SceneData(prims, bvs, float(hs.globalRayDensity), bool(hs.useBoundingVolumes))


def load_xml(path: str | Path) -> SceneData:
    """Scene::readScene (Scene.cpp:24-52) through the native loader."""
    hs = HostScene()
    rc = lib().rf_scene_load_xml(str(path).encode(), byref(hs))
    if rc != 0:
        raise RayFactorError(rc, f"rf_scene_load_xml({path})")
    try:
        return _from_host(hs)
    finally:
        lib().rf_host_scene_free(byref(hs))


def parse_xml(text: str) -> SceneData:
    raw = text.encode()
    hs = HostScene()
    rc = lib().rf_scene_parse_xml(raw, len(raw), byref(hs))
    if rc != 0:
        raise RayFactorError(rc, "rf_scene_parse_xml")
    try:
        return _from_host(hs)
    finally:
        lib().rf_host_scene_free(byref(hs))


def read_rfscene(path: str | Path) -> SceneData:
    """Scene bytes written by the reference loader's dump driver (fixtures)."""
    b = Path(path).read_bytes()
    if b[:4] != b"RFSC":
        raise ValueError(f"{path}: not an rfscene file")
    ver, n, nb = np.frombuffer(b[4:16], dtype="<i4")
    grd = float(np.frombuffer(b[16:20], dtype="<f4")[0])
    use_bv = bool(np.frombuffer(b[20:24], dtype="<i4")[0])
    if ver != 1:
        raise ValueError(f"{path}: unsupported version {ver}")
    off = 24
    prims = np.frombuffer(b[off:off + 112 * n], dtype=PRIM_DTYPE).copy()
    off += 112 * n
    bvs = np.frombuffer(b[off:off + 72 * nb], dtype=BV_DTYPE).copy()
    return SceneData(prims, bvs, grd, use_bv)


# ---------------------------------------------------------------------------
# XML writer helpers
# ---------------------------------------------------------------------------
def _f(x: float) -> str:
    return repr(float(x))


def _prim(kind: str, *, bounding: bool | None = None, small: float | None = None, large: float | None = None,
          scale=None, rotate=None, translate=None, indent: str = "\t\t") -> str:
    attrs = f' type="{kind}"'
    if bounding is not None:
        attrs += f' isBounding="{"true" if bounding else "false"}"'
    body = []
    if small is not None:
        body.append(f'<smallRadius value="{_f(small)}" />')
    if large is not None:
        body.append(f'<largeRadius value="{_f(large)}" />')
    if scale is not None:
        body.append(f'<scale x="{_f(scale[0])}" y="{_f(scale[1])}" z="{_f(scale[2])}" />')
    if rotate is not None:
        deg, ax = rotate
        body.append(f'<rotate degrees="{_f(deg)}" x="{_f(ax[0])}" y="{_f(ax[1])}" z="{_f(ax[2])}" />')
    if translate is not None:
        body.append(f'<translation x="{_f(translate[0])}" y="{_f(translate[1])}" z="{_f(translate[2])}" />')
    inner = "".join(f"\n{indent}\t{b}" for b in body)
    return f"{indent}<primitive{attrs} >{inner}\n{indent}</primitive>\n"


def _doc(description: str, body: str, *, use_bv: bool = False, ray_density: float = 1000000) -> str:
    bv = '\n\t\t<useBoundingVolumes value="true" />' if use_bv else ""
    return ('<?xml version="1.0" ?>\n<project>\n\t<settings>\n'
            f'\t\t<globalRayDensity value="{_f(ray_density)}" />\n'
            f'\t\t<description>{description}</description>{bv}\n'
            f'\t</settings>\n\t<geometry>\n{body}\t</geometry>\n</project>\n')


FLIP_X = (180.0, (1.0, 0.0, 0.0))   # normal +z -> -z


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------
def config1_xml() -> str:
    """Closed cylinder enclosure r=1, H=2 (inward wall + two end discs) and a
    coaxial pair of unit discs at spacing 1 placed at x = 10."""
    body = (_prim("cylinderSurface", bounding=True, small=1.0, scale=(1, 1, 2))
            + _prim("disc", scale=(1, 1, 1))
            + _prim("disc", scale=(1, 1, 1), rotate=FLIP_X, translate=(0, 0, 2))
            + _prim("disc", scale=(1, 1, 1), translate=(10, 0, 0))
            + _prim("disc", scale=(1, 1, 1), rotate=FLIP_X, translate=(10, 0, 1)))
    return _doc("closed cylinder enclosure and coaxial discs", body)


def _furnace_parts(wall_bands: int, wall_r: float, height: float, preform_r: float, preform_bands: int,
                   preform_z0: float, neck_deg: float | None):
    """Primitive groups of a coaxial furnace: wall bands, preform bands, optional
    conical neck below the preform tip with its end disc (else a disc closing the
    preform), bottom and top annuli between preform and wall radius."""
    groups = {}
    hw = height / wall_bands
    groups["wall"] = [_prim("cylinderSurface", bounding=True, small=1.0, scale=(wall_r, wall_r, hw),
                            translate=(0, 0, i * hw)) for i in range(wall_bands)]
    hp = (height - preform_z0) / preform_bands
    groups["preform"] = [_prim("cylinderSurface", bounding=False, small=1.0, scale=(preform_r, preform_r, hp),
                               translate=(0, 0, preform_z0 + i * hp)) for i in range(preform_bands)]
    if neck_deg is not None:
        r_tip = 0.3 * preform_r
        length = (preform_r - r_tip) / math.tan(math.radians(neck_deg))
        z_tip = preform_z0 - length
        groups["neck"] = [_prim("frustum", bounding=False, small=preform_r / r_tip, scale=(r_tip, r_tip, length),
                                translate=(0, 0, z_tip))]
        groups["end"] = [_prim("disc", scale=(r_tip, r_tip, 1), rotate=FLIP_X, translate=(0, 0, z_tip))]
    else:
        groups["end"] = [_prim("disc", scale=(preform_r, preform_r, 1), rotate=FLIP_X, translate=(0, 0, preform_z0))]
    groups["annuli"] = [
        _prim("annulus", small=preform_r, large=wall_r),
        _prim("annulus", small=preform_r, large=wall_r, rotate=FLIP_X, translate=(0, 0, height)),
    ]
    return groups


def furnace_xml(with_bv: bool = False, wall_bands: int = 180, wall_r: float = 0.5, height: float = 3.0,
                preform_r: float = 0.1, preform_bands: int = 60, neck_deg: float | None = 20.0) -> str:
    """Furnace-like scene: 180 inward wall bands (r 0.5, z in [0, 3]), a 60-band
    preform (r 0.1, z in [1, 3]), a 20-degree conical neck below it closed by an
    end disc, and bottom / top annuli (r 0.1-0.5): 244 primitives.  With
    with_bv, every group sits under a bounding surface of its own shape (one
    bounding cylinder over the wall bands, one over the preform bands)."""
    g = _furnace_parts(wall_bands, wall_r, height, preform_r, preform_bands, height / 3.0, neck_deg)
    order = ["wall", "preform", "neck", "end", "annuli"]
    if not with_bv:
        body = "".join("".join(g[k]) for k in order if k in g)
        return _doc("furnace: wall bands, preform bands, conical neck, end disc, annuli", body)

    def bv(kind, children, scale=None, translate=None, small=None, rotate=None, large=None):
        inner = ""
        if small is not None:
            inner += f'\n\t\t\t<smallRadius value="{_f(small)}" />'
        if large is not None:
            inner += f'\n\t\t\t<largeRadius value="{_f(large)}" />'
        if scale is not None:
            inner += f'\n\t\t\t<scale x="{_f(scale[0])}" y="{_f(scale[1])}" z="{_f(scale[2])}" />'
        if rotate is not None:
            deg, ax = rotate
            inner += f'\n\t\t\t<rotate degrees="{_f(deg)}" x="{_f(ax[0])}" y="{_f(ax[1])}" z="{_f(ax[2])}" />'
        if translate is not None:
            inner += f'\n\t\t\t<translation x="{_f(translate[0])}" y="{_f(translate[1])}" z="{_f(translate[2])}" />'
        kids = "".join(c.replace("\t\t<", "\t\t\t\t<").replace("\t\t</", "\t\t\t\t</") for c in children)
        return f'\t\t<bv type="{kind}">{inner}\n\t\t\t<children>\n{kids}\t\t\t</children>\n\t\t</bv>\n'

    z0 = height / 3.0
    body = bv("cylinderSurface", g["wall"], scale=(wall_r, wall_r, height))
    body += bv("cylinderSurface", g["preform"], scale=(preform_r, preform_r, height - z0), translate=(0, 0, z0))
    for k in ("neck", "end"):
        for c in g.get(k, []):
            body += _single_bv(c)
    for c in g["annuli"]:
        body += _single_bv(c)
    return _doc("furnace with bounding surfaces", body, use_bv=True)


def _single_bv(prim_xml: str) -> str:
    """A bounding surface identical to its one child primitive."""
    head, rest = prim_xml.split(">", 1)
    inner, _ = rest.rsplit("</primitive>", 1)
    kind = head.split('type="')[1].split('"')[0]
    lines = [l.strip() for l in inner.strip().splitlines() if l.strip()]
    body = "".join(f"\n\t\t\t{l}" for l in lines)
    child = prim_xml.replace("\t\t<", "\t\t\t\t<").replace("\t\t</", "\t\t\t\t</")
    return f'\t\t<bv type="{kind}">{body}\n\t\t\t<children>\n{child}\t\t\t</children>\n\t\t</bv>\n'


def _axis_rotation(d):
    """(degrees, axis) rotating +z onto unit vector d."""
    d = np.asarray(d, dtype=np.float64)
    z = np.array([0.0, 0.0, 1.0])
    c = float(np.clip(d @ z, -1.0, 1.0))
    ax = np.cross(z, d)
    s = np.linalg.norm(ax)
    if s < 1e-12:
        return (0.0, (1.0, 0.0, 0.0)) if c > 0 else (180.0, (1.0, 0.0, 0.0))
    ax = ax / s
    return (math.degrees(math.acos(c)), tuple(ax))


def soup_xml(n: int = 1000, seed: int = 0x52465331) -> str:
    """Random primitive soup: types uniform over {cylinder, frustum, disc,
    annulus, sphere}; centres U[0,10]^3; radii log-U[0.05, 0.5]; lengths
    U[0.1, 1]; axes uniform on the sphere; outward normals."""
    rng = np.random.default_rng(seed)
    kinds = ["cylinderSurface", "frustum", "disc", "annulus", "sphere"]
    out = []
    for _ in range(n):
        k = kinds[rng.integers(0, 5)]
        c = rng.uniform(0.0, 10.0, 3)
        r = float(np.exp(rng.uniform(math.log(0.05), math.log(0.5))))
        length = float(rng.uniform(0.1, 1.0))
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        rot = _axis_rotation(v)
        if k == "cylinderSurface":
            out.append(_prim(k, bounding=False, small=1.0, scale=(r, r, length), rotate=rot, translate=c))
        elif k == "frustum":
            ratio = float(rng.uniform(0.0, 1.0))
            out.append(_prim(k, bounding=False, small=ratio, scale=(r, r, length), rotate=rot, translate=c))
        elif k == "disc":
            out.append(_prim(k, scale=(r, r, 1.0), rotate=rot, translate=c))
        elif k == "annulus":
            inner = r * float(rng.uniform(0.1, 0.9))
            out.append(_prim(k, small=inner, large=r, rotate=rot, translate=c))
        else:
            out.append(_prim(k, bounding=False, scale=(r, r, r), translate=c))
    return _doc(f"random primitive soup n={n} seed={seed}", "".join(out))


def sweep_params(n_geoms: int = 256, seed: int = 0x52465331):
    """Per-geometry parameters of the batched design sweep."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_geoms):
        out.append(dict(
            wall_bands=int(rng.integers(20, 501)), wall_r=float(rng.uniform(0.3, 0.8)),
            height=float(rng.uniform(1.0, 4.0)), preform_r=float(rng.uniform(0.05, 0.2)),
            preform_bands=int(rng.integers(10, 101)),
            neck_deg=float(rng.uniform(10.0, 40.0)) if rng.uniform() < 0.5 else None))
    return out


def sweep_xml(params: dict) -> str:
    g = _furnace_parts(params["wall_bands"], params["wall_r"], params["height"], params["preform_r"],
                       params["preform_bands"], params["height"] / 3.0, params["neck_deg"])
    body = "".join("".join(g[k]) for k in ["wall", "preform", "neck", "end", "annuli"] if k in g)
    return _doc("furnace design-sweep geometry", body)


def sweep(n_geoms: int = 256, seed: int = 0x52465331) -> list[SceneData]:
    """Batched design sweep: coaxial furnaces with wall bands U{20..500}, wall
    radius U[0.3, 0.8], height U[1, 4], preform radius U[0.05, 0.2] with
    U{10..100} bands, and a conical neck (half-angle U[10, 40] degrees) with
    probability 1/2."""
    return [parse_xml(sweep_xml(p)) for p in sweep_params(n_geoms, seed)]


def streaming_requests(n: int, seed: int = 0x52465331, rate_hz: float = 2.0, max_prims: int = 2000,
                       min_rays: float = 1e5, max_rays: float = 1e8):
    """Poisson-arrival solve requests: (arrival_s, scene, rays_per_emitter, kind).
    Scene sizes U{10..max_prims} drawn as a furnace or a soup (kind); budgets
    log-U[min_rays, max_rays] (default [1e5, 1e8])."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate_hz))
        size = int(rng.integers(10, max_prims + 1))
        if rng.uniform() < 0.5:
            wb = max(1, int(size * 0.75) - 3)
            pb = max(1, size - wb - 4)
            sc, kind = parse_xml(furnace_xml(wall_bands=wb, preform_bands=pb)), "furnace"
        else:
            sc, kind = parse_xml(soup_xml(size, seed=seed + i + 1)), "soup"
        rays = int(round(float(np.exp(rng.uniform(math.log(min_rays), math.log(max_rays))))))
        out.append((t, sc, rays, kind))
    return out
