"""Pinned transcendental and the hit tests of the oracle, checked against
mathematics: libm double sin/cos error bounds, analytic intersections of
constructed rays, stable-root identities and the reference quirks."""
import math

import numpy as np
import pytest

import oracle_pins
from conftest import GOLDEN

INF = float("inf")
CYL, FRU, SPH, ANN, REC, DISC = 0, 1, 2, 3, 4, 5
I12 = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]


def ulp32(x):
    return np.spacing(np.float32(abs(x)))


def test_sincosf_error_bound(oracle):
    # the kernel's arguments are u * 2pi_f with u in [0, 1): sweep that domain densely
    u = np.linspace(0.0, 1.0, 20001, endpoint=False).astype(np.float32)
    two_pi = np.float32(2.0) * np.float32(np.pi)
    worst = 0.0
    for x in (u * two_pi):
        s, c = oracle.sincosf(float(x))
        ts, tc = math.sin(float(x)), math.cos(float(x))
        # absolute error bounded by 2 ulp of 1 (results of magnitude <= 1)
        worst = max(worst, abs(s - ts), abs(c - tc))
        assert abs(s - ts) <= 2 * ulp32(max(abs(ts), 1e-30)) + 2 ** -26
        assert abs(c - tc) <= 2 * ulp32(max(abs(tc), 1e-30)) + 2 ** -26
    assert worst < 1.2e-7


def test_sincosf_quadrants_exact_points(oracle):
    assert oracle.sincosf(0.0) == (0.0, 1.0)
    s, c = oracle.sincosf(float(np.float32(np.pi / 2)))
    assert c == pytest.approx(0.0, abs=1e-7) and s == 1.0


def test_sincos64_error_bound(oracle):
    import mpmath as mp
    mp.mp.prec = 120
    for x in np.linspace(0, 2 * np.pi, 997):
        s, c = oracle.sincos64(float(x))
        ts, tc = float(mp.sin(mp.mpf(float(x)))), float(mp.cos(mp.mpf(float(x))))
        assert abs(s - ts) <= 2 * np.spacing(max(abs(ts), 1e-300)) + 2 ** -54
        assert abs(c - tc) <= 2 * np.spacing(max(abs(tc), 1e-300)) + 2 ** -54


def ray(s, c):
    return np.array(list(s) + list(c), dtype=np.float64)


@pytest.mark.parametrize("mode", ["f32", "f64"])
def test_cylinder_from_axis(oracle, mode):
    hit = oracle.hit_f32 if mode == "f32" else oracle.hit_f64
    # unit cylinder z in [0,1]: from the axis at z=0.5 along +x hits at t=1
    assert hit(CYL, I12, 0, 1, ray((0, 0, 0.5), (1, 0, 0))) == pytest.approx(1.0, rel=1e-6)
    # oblique: direction (3,4,0)/5 hits at t=1
    assert hit(CYL, I12, 0, 1, ray((0, 0, 0.5), (0.6, 0.8, 0))) == pytest.approx(1.0, rel=1e-6)
    # outside, heading in: from x=-3 enters at t=2
    assert hit(CYL, I12, 0, 1, ray((-3, 0, 0.5), (1, 0, 0))) == pytest.approx(2.0, rel=1e-6)
    # misses above the finite cylinder
    assert hit(CYL, I12, 0, 1, ray((0, 0, 0.5), (0.6, 0, 0.8))) == INF
    # passes the infinite cylinder's axis-parallel line: no root
    assert hit(CYL, I12, 0, 1, ray((-3, 2, 0.5), (1, 0, 0))) == INF or math.isnan(
        hit(CYL, I12, 0, 1, ray((-3, 2, 0.5), (1, 0, 0))))


@pytest.mark.parametrize("mode", ["f32", "f64"])
def test_frustum_cone(oracle, mode):
    hit = oracle.hit_f32 if mode == "f32" else oracle.hit_f64
    # cone radius 1 - z (util1 = 0): from the axis at z=0.5 along +x hits at t = 0.5
    assert hit(FRU, I12, 0.0, 1.0, ray((0, 0, 0.5), (1, 0, 0))) == pytest.approx(0.5, rel=1e-6)
    # frustum radius 1 + z (util1 = 2) at z = 0.25 -> 1.25
    assert hit(FRU, I12, 2.0, 1.0, ray((0, 0, 0.25), (0, 1, 0))) == pytest.approx(1.25, rel=1e-6)


def test_sphere_quirk_q1(oracle):
    # unit sphere from x=-3 along +x: true entry at t = 2
    r = ray((-3, 0, 0), (1, 0, 0))
    assert oracle.hit_f64(SPH, I12, 0, 1, r) == pytest.approx(2.0, rel=1e-12)
    # reference formula omits the square root (RayFactorGPU.cl:272-278):
    # B = -3, A = 1, C = 8, discrim = 9 - 8 = 1 -> t1 = (3 - 1) = 2 (coincides here)
    assert oracle.hit_f32(SPH, I12, 0, 1, r) == pytest.approx(2.0, rel=1e-6)
    # from x=-5: true t = 4; reference discrim = 25 - 24 = 1 -> t = 5 - 1 = 4 coincides again
    # a glancing ray separates them: s = (-3, 0.6, 0): C = 8.36, B = -3, D = 0.64
    r = ray((-3, 0.6, 0), (1, 0, 0))
    assert oracle.hit_f64(SPH, I12, 0, 1, r) == pytest.approx(3.0 - 0.8, rel=1e-12)
    assert oracle.hit_f32(SPH, I12, 0, 1, r) == pytest.approx(3.0 - 0.64, rel=1e-6)


@pytest.mark.parametrize("mode", ["f32", "f64"])
def test_planes(oracle, mode):
    hit = oracle.hit_f32 if mode == "f32" else oracle.hit_f64
    down = ray((0.5, 0.5, 2), (0, 0, -1))
    assert hit(DISC, I12, 0, 0, down) == pytest.approx(2.0)
    assert hit(REC, I12, 0, 0, down) == pytest.approx(2.0)
    assert hit(ANN, I12, 0.25, 1.0, down) == pytest.approx(2.0)     # r^2 = 0.5 in [0.25, 1]
    assert hit(ANN, I12, 0.6, 1.0, down) == INF
    out = ray((1.5, 0.0, 2), (0, 0, -1))
    assert hit(DISC, I12, 0, 0, out) == INF
    assert hit(REC, I12, 0, 0, ray((0.9, 0.9, 2), (0, 0, -1))) == pytest.approx(2.0)
    assert hit(DISC, I12, 0, 0, ray((0.9, 0.9, 2), (0, 0, -1))) == INF
    # behind the origin
    assert hit(DISC, I12, 0, 0, ray((0, 0, 2), (0, 0, 1))) == INF


def test_eps_acceptance_quirk_q7(oracle):
    # rectangle accepts t == EPS, disc does not (RayFactorGPU.cl:137 vs :158)
    e = float(np.float32(1e-3))
    r = ray((0, 0, e), (0, 0, -1))
    assert oracle.hit_f32(REC, I12, 0, 0, r) == pytest.approx(e)
    assert oracle.hit_f32(DISC, I12, 0, 0, r) == INF


def test_stable_roots_agree_with_textbook(oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        s = rng.uniform(-3, 3, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        t = oracle.hit_f64(SPH, I12, 0, 1, ray(s, d))
        A, B, C = d @ d, s @ d, s @ s - 1
        D = B * B - A * C
        roots = [] if D < 0 else sorted([(-B - math.sqrt(D)) / A, (-B + math.sqrt(D)) / A])
        good = [r for r in roots if r > 1e-3]
        assert (t == INF) if not good else t == pytest.approx(good[0], rel=1e-9, abs=1e-12)


@pytest.mark.parametrize("mode", ["compat", "fp64"])
def test_find_object_tie_lowest_index(oracle, mode):
    oracle_pins.pin_find_tie(oracle, mode)


@pytest.mark.parametrize("mode", ["compat", "fp64"])
def test_bv_stale_index_quirk_q12(oracle, mode):
    """runBVGPU: an intermediate-check BV sets objectIndex; a later plain BV that
    wins keeps that stale index (RayFactorGPU.cl:596-619); control without it."""
    oracle_pins.pin_bv_stale_index(oracle, mode)


@pytest.mark.parametrize("mode", ["compat", "fp64"])
def test_bv_intermediate_check_miss(oracle, mode):
    oracle_pins.pin_bv_intermediate_miss(oracle, mode)
