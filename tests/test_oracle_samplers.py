"""Emission samplers of the oracle (RayFactorGPU.cl:298-443 in compat form, the
FP64 physics form): points lie on the surface, directions are unit vectors on
the emitting side, positions are area-uniform and directions cosine-weighted."""
import math

import numpy as np
import pytest
from scipy import stats

import rayfactorcl_amd as rf
from rayfactorcl_amd.scenes import _doc, _prim

SHAPES = {
    "disc": _prim("disc", scale=(0.7, 0.7, 1), rotate=(30, (0.6, 0.8, 0)), translate=(1, 2, 3)),
    "annulus": _prim("annulus", small=0.4, large=1.3, rotate=(120, (0, 0.6, 0.8)), translate=(-1, 0, 2)),
    "cylinder_out": _prim("cylinderSurface", bounding=False, small=1.0, scale=(0.5, 0.5, 2.0),
                          rotate=(45, (1, 0, 0)), translate=(0, 1, 0)),
    "cylinder_in": _prim("cylinderSurface", bounding=True, small=1.0, scale=(1.5, 1.5, 0.5)),
    "frustum": _prim("frustum", bounding=False, small=0.4, scale=(1.0, 1.0, 1.5), rotate=(60, (0, 1, 0)),
                     translate=(2, 0, 0)),
    "cone_in": _prim("cylinderSurface", bounding=True, small=0.0),
    "rectangle": _prim("rectangle", scale=(2.0, 0.5, 1), rotate=(90, (1, 0, 0)), translate=(0, -1, 1)),
    "sphere": _prim("sphere", bounding=False, scale=(0.8, 0.8, 0.8), translate=(1, 1, 1)),
    "sphere_in": _prim("sphere", bounding=True, scale=(2, 2, 2)),
}


def shape(name):
    return rf.parse_xml(_doc("sampler test", SHAPES[name])).prims[0]


def world_normal(p, s):
    """Unit world normal at world point s (object-space gradient through invm^T), with util2 flips."""
    iv = np.asarray(p["invm"], dtype=np.float64).reshape(3, 4)
    q = iv[:, :3] @ s + iv[:, 3]
    t = int(p["type"])
    if t in (rf.DISC, rf.ANNULUS, rf.RECTANGLE):
        g = np.array([0.0, 0.0, 1.0])
    elif t == rf.CYLINDER:
        g = np.array([q[0], q[1], 0.0]) * p["util2"]
    elif t == rf.FRUSTUM:
        u1 = float(p["util1"])
        g = np.array([q[0], q[1], (1 - u1) * (1 + (u1 - 1) * q[2])]) * p["util2"]
    else:
        g = q * p["util2"]
    n = iv[:, :3].T @ g
    return n / np.linalg.norm(n), q


def residual(p, q):
    t = int(p["type"])
    r2 = q[0] ** 2 + q[1] ** 2
    if t == rf.DISC:
        return abs(q[2]) + max(0.0, r2 - 1)
    if t == rf.ANNULUS:
        return abs(q[2]) + max(0.0, r2 - p["util2"]) + max(0.0, p["util1"] - r2)
    if t == rf.RECTANGLE:
        return abs(q[2]) + max(0.0, abs(q[0]) - 1) + max(0.0, abs(q[1]) - 1)
    if t == rf.CYLINDER:
        return abs(math.sqrt(r2) - 1) + max(0.0, -q[2]) + max(0.0, q[2] - 1)
    if t == rf.FRUSTUM:
        return abs(math.sqrt(r2) - (1 + (p["util1"] - 1) * q[2])) + max(0.0, -q[2]) + max(0.0, q[2] - 1)
    return abs(math.sqrt(r2 + q[2] ** 2) - 1)


@pytest.mark.parametrize("name", list(SHAPES))
@pytest.mark.parametrize("mode", ["f32", "f64"])
def test_sampler_invariants(oracle, name, mode):
    p = shape(name)
    rng = np.random.default_rng(7)
    emit = oracle.emit_f32 if mode == "f32" else oracle.emit_f64
    tol = 2e-5 if mode == "f32" else 1e-6
    cos_t = []
    for _ in range(400):
        u = (rng.integers(0, 2 ** 24, 4) * 2.0 ** -24)
        r = emit(p, u.astype(np.float32) if mode == "f32" else u)
        s, c = np.asarray(r[:3], dtype=np.float64), np.asarray(r[3:], dtype=np.float64)
        n, q = world_normal(p, s)
        # m and invm are float32 inverses of each other: 1e-6 relative residual
        assert residual(p, q) < 2e-5, (name, q)
        assert abs(np.linalg.norm(c) - 1) < tol
        cos_t.append(c @ n)
    assert min(cos_t) > -tol


@pytest.mark.parametrize("name", ["disc", "cylinder_out", "frustum", "sphere", "annulus"])
def test_cosine_weighted_and_area_uniform(oracle, name):
    p = shape(name)
    rng = np.random.default_rng(11)
    N = 4000
    cos2, area_stat = [], []
    u1 = float(p["util1"])
    for _ in range(N):
        u = rng.integers(0, 2 ** 24, 4) * 2.0 ** -24
        r = oracle.emit_f64(p, u)
        s, c = np.asarray(r[:3]), np.asarray(r[3:])
        n, q = world_normal(p, s)
        cos2.append((c @ n) ** 2)
        t = int(p["type"])
        if t == rf.DISC:
            area_stat.append(q[0] ** 2 + q[1] ** 2)              # r^2 ~ U[0,1]
        elif t == rf.ANNULUS:
            area_stat.append((q[0] ** 2 + q[1] ** 2 - u1) / (p["util2"] - u1))
        elif t == rf.CYLINDER:
            area_stat.append(q[2])                                # z ~ U[0,1]
        elif t == rf.FRUSTUM:
            z = q[2]                                              # CDF of radius-weighted z
            area_stat.append((z + (u1 - 1) * z * z / 2) / (1 + (u1 - 1) / 2))
        else:
            area_stat.append((q[2] + 1) / 2)                      # Archimedes: z ~ U[-1,1]
    # cosine-weighted hemisphere: cos^2(theta) ~ U[0,1]
    assert stats.kstest(cos2, "uniform").pvalue > 1e-3
    assert stats.kstest(area_stat, "uniform").pvalue > 1e-3
