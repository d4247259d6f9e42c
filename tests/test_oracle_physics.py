"""Whole-solve checks of the oracle against what the mathematics fixes:
Howell-catalogue closed forms (sample scenes, their bounding-surface variants
and the config-1 scene), reciprocity and closure.

Each measured view factor must equal the closed form less the first-order
seam loss that the EPS = 1e-3 self-intersection cut-off causes at touching
edges (derived in closed_forms.py), within 3 standard errors computed from
per-work-item hit counts (the 4 rays of a work item share one random block)."""
import math

import numpy as np
import pytest

from closed_forms import CASES, K, c122_quadrature, check, seam_loss, vf_and_sigma

N_WI = 1 << 19            # 2^21 rays per emitter: sigma ~ 3e-4 at F = 0.2, 3e-5 at F = 0.999


def measure(oracle, sc, i, j, mode):
    idx = oracle.trace(sc.prims, sc.bvs, i, 0, N_WI, mode=mode, use_bvs=sc.useBoundingVolumes)
    return vf_and_sigma(idx, j)


def test_clamped_cosine_closed_form():
    """K(b) = E[max(0, d.v)] over cosine-weighted d, against 2-D quadrature."""
    from scipy import integrate
    for b in (0.0, math.pi / 6, math.pi / 4, math.pi / 2, 2.0, math.pi):
        v = (-math.sin(b), 0.0, math.cos(b))
        f = lambda ph, th: max(0.0, v[0] * math.sin(th) * math.cos(ph) + v[2] * math.cos(th)) * \
            math.cos(th) * math.sin(th) / math.pi
        q, _ = integrate.dblquad(f, 0.0, math.pi / 2, 0.0, 2 * math.pi, epsabs=1e-10)
        assert q == pytest.approx(K(b), abs=1e-8), b
    assert K(math.pi / 2) == pytest.approx(2 / (3 * math.pi))


@pytest.mark.parametrize("mode", ["compat", "fp64"])
@pytest.mark.parametrize("scene,i,j,exact,loss", CASES)
def test_closed_forms(oracle, golden_scene, mode, scene, i, j, exact, loss):
    sc = golden_scene(scene)
    assert sc.useBoundingVolumes == scene.endswith("-BV")
    vf, se = measure(oracle, sc, i, j, mode)
    assert check(vf, se, exact, loss), (vf, exact - loss, se)


def test_seam_loss_is_resolved(oracle, golden_scene):
    """C-109 disc -> cone: the predicted loss (1.42e-3) is ~50 standard errors;
    the closed form without it fails, with it passes (both forms)."""
    sc = golden_scene("input-C109")
    _, _, _, exact, loss = next(c for c in CASES if c[0] == "input-C109" and c[1] == 1)
    for mode in ("compat", "fp64"):
        vf, se = measure(oracle, sc, 1, 0, mode)
        assert loss > 40 * se
        assert check(vf, se, exact, loss) and not check(vf, se, exact, 0.0)


def test_c122_sphere_to_rectangle(oracle, golden_scene):
    """Howell C-122 (sphere over the corner of a rectangle), both directions.
    The reference's sphere test without the square root (quirk Q1) keeps the
    sign of the discriminant, so it classifies hits and misses correctly and
    only misplaces t: with nothing else in the scene the compat form also
    matches (Q1 itself is pinned at the hit level in test_oracle_math)."""
    sc = golden_scene("input-C122")
    exact = c122_quadrature()
    assert exact == pytest.approx(math.atan(16 / 9 / math.sqrt(41 / 9)) / (4 * math.pi), rel=1e-6)
    vf, se = measure(oracle, sc, 0, 1, "fp64")
    assert check(vf, se, exact, 0.0)
    vf, se = measure(oracle, sc, 1, 0, "fp64")
    assert check(vf, se, math.pi * exact, 0.0)                   # reciprocity, A_s / A_r = pi
    vfc, sec = measure(oracle, sc, 1, 0, "compat")
    assert check(vfc, sec, math.pi * exact, 0.0)


def test_reciprocity_and_closure(oracle, golden_scene):
    sc = golden_scene("config1")
    hits, rays = oracle.solve(sc.prims, None, n_wi=1 << 16, mode="fp64", emitters=(0, 3))
    F = hits[:3, :3] / rays[:3, None]
    A = np.array([4 * math.pi, math.pi, math.pi])
    for a in range(3):
        for b in range(3):
            assert A[a] * F[a, b] == pytest.approx(A[b] * F[b, a], rel=0.02, abs=2e-3)
    # closure: each row of the closed enclosure sums to 1 less the rays lost
    # at its rims (wall: two rims, end disc: one), i.e. the fraction of rays
    # that hit nothing is that seam loss
    lost = [2 * seam_loss(2 * math.pi, 4 * math.pi, math.pi / 2), seam_loss(2 * math.pi, math.pi, math.pi / 2),
            seam_loss(2 * math.pi, math.pi, math.pi / 2)]
    for e in range(3):
        miss, se = measure(oracle, sc, e, -1, "fp64")
        assert check(miss, se, lost[e], 0.0), (e, miss, lost[e], se)
        assert miss > 0.0


def test_compat_and_fp64_use_the_same_stream(oracle, golden_scene):
    # same rays up to the emission frame: the disc -> disc pair (no frame quirk
    # in play for an unrotated emitter) agrees ray for ray on most rays
    sc = golden_scene("input-C47")
    a = oracle.trace(sc.prims, None, 0, 0, 4096, mode="compat")
    b = oracle.trace(sc.prims, None, 0, 0, 4096, mode="fp64")
    assert (a == b).mean() > 0.99
