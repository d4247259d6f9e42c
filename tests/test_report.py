"""Surface areas and the closure / reciprocity report (CPU).

Areas are pinned by closed forms that do not share the library's quadrature:
lateral areas of circular and elliptic cylinders and cones, spheroid and
scalene-ellipsoid areas (Legendre's elliptic-integral form via scipy), annulus
and rectangle areas, and invariance under rotation.  The report is pinned by
the exact view-factor matrix of the closed-cylinder configuration
(coaxial-disc closed form, summation and reciprocity rules), whose closure is
1 for the enclosure rows and whose reciprocity error is zero.
"""
import math

import numpy as np
import pytest
from scipy import special

import rayfactorcl_amd as rf
from rayfactorcl_amd import scenes


def _one(kind, **kw):
    body = scenes._prim(kind, **kw)
    return scenes.parse_xml(scenes._doc("one primitive", body)).prims


def _area(kind, **kw):
    return float(rf.areas(_one(kind, **kw))[0])


def _ellipse_perimeter(a, b):
    a, b = max(a, b), min(a, b)
    return 4 * a * special.ellipe(1 - (b / a) ** 2)


def _ellipsoid_area(a, b, c):
    a, b, c = sorted((a, b, c), reverse=True)
    if a == c:
        return 4 * math.pi * a * a
    phi = math.acos(c / a)
    m = (a * a * (b * b - c * c)) / (b * b * (a * a - c * c))
    E = special.ellipeinc(phi, m)
    F = special.ellipkinc(phi, m)
    return 2 * math.pi * c * c + 2 * math.pi * a * b / math.sin(phi) * (E * math.sin(phi) ** 2 + F * math.cos(phi) ** 2)


def test_config1_areas():
    prims = scenes.parse_xml(scenes.config1_xml()).prims
    np.testing.assert_allclose(rf.areas(prims), [4 * math.pi] + [math.pi] * 4, rtol=1e-12)


@pytest.mark.parametrize("kind,kw,exact", [
    ("cylinderSurface", dict(scale=(1.5, 1.5, 3.0)), 2 * math.pi * 1.5 * 3.0),
    ("cylinderSurface", dict(scale=(2.0, 0.5, 1.25)), _ellipse_perimeter(2.0, 0.5) * 1.25),
    # radius 2 at z = 0, 2 * 0.5 at the top, height 3: pi (r1 + r2) * slant
    ("frustum", dict(small=0.5, scale=(2.0, 2.0, 3.0)), math.pi * 3.0 * math.sqrt(1.0 + 9.0)),
    ("frustum", dict(small=1.75, scale=(0.5, 0.5, 0.2), rotate=(33.0, (1 / 3, 2 / 3, 2 / 3))),
     math.pi * (0.5 + 0.875) * math.hypot(0.375, 0.2)),
    ("sphere", dict(scale=(0.7, 0.7, 0.7)), 4 * math.pi * 0.49),
    ("sphere", dict(scale=(1.0, 1.0, 2.0)), _ellipsoid_area(1.0, 1.0, 2.0)),
    ("sphere", dict(scale=(2.0, 2.0, 1.0)), _ellipsoid_area(2.0, 2.0, 1.0)),
    ("sphere", dict(scale=(3.0, 2.0, 1.0), rotate=(71.0, (0.6, 0.0, -0.8))), _ellipsoid_area(3.0, 2.0, 1.0)),
    ("disc", dict(scale=(2.0, 0.5, 1.0)), math.pi * 1.0),
    ("annulus", dict(small=0.5, large=2.0), math.pi * (4.0 - 0.25)),
    ("annulus", dict(small=0.25, large=1.0, scale=(3.0, 3.0, 1.0)), 9 * math.pi * (1.0 - 0.0625)),
    ("rectangle", dict(scale=(2.0, 3.0, 1.0), rotate=(40.0, (0.8, 0.6, 0.0))), 24.0),
])
def test_areas_closed_forms(kind, kw, exact):
    # float32 transforms: the rotation's matrix entries carry ~1e-7 error
    assert _area(kind, **kw) == pytest.approx(exact, rel=2e-6)


def test_areas_rotation_invariant():
    # the loader, like the reference, does not normalise the rotation axis:
    # unit axes give rotations
    base = _area("frustum", small=0.3, scale=(1.0, 2.0, 0.5))
    for deg, ax in [(17.0, (1.0, 0.0, 0.0)), (123.0, (2 / 3, -1 / 3, 2 / 3)), (270.0, (0.0, 0.0, 1.0))]:
        assert _area("frustum", small=0.3, scale=(1.0, 2.0, 0.5), rotate=(deg, ax)) == pytest.approx(base, rel=2e-6)


def test_areas_errors():
    prims = scenes.parse_xml(scenes.config1_xml()).prims.copy()
    prims["type"][1] = 99
    with pytest.raises(rf.RayFactorError):
        rf.areas(prims)
    assert len(rf.areas(prims[:0])) == 0


def _config1_exact():
    """Exact view factors of config1's five primitives (wall, bottom disc,
    top disc, then two coaxial unit discs at spacing 1)."""
    def disc_pair(r, h):
        R = r / h
        X = 1 + (1 + R * R) / (R * R)
        return 0.5 * (X - math.sqrt(X * X - 4))
    f_dd = disc_pair(1.0, 2.0)                 # end disc -> end disc
    f_dw = 1 - f_dd                            # end disc -> wall (summation)
    f_wd = math.pi * f_dw / (4 * math.pi)      # reciprocity
    f_ww = 1 - 2 * f_wd
    f_pp = disc_pair(1.0, 1.0)
    vf = np.zeros((5, 5))
    vf[0, 0], vf[0, 1], vf[0, 2] = f_ww, f_wd, f_wd
    vf[1, 0], vf[1, 2] = f_dw, f_dd
    vf[2, 0], vf[2, 1] = f_dw, f_dd
    vf[3, 4] = vf[4, 3] = f_pp
    return vf, f_pp


def test_report_exact_matrix():
    prims = scenes.parse_xml(scenes.config1_xml()).prims
    vf, f_pp = _config1_exact()
    rep = rf.vf_report(vf, rf.areas(prims))
    np.testing.assert_allclose(rep["closure"], [1, 1, 1, f_pp, f_pp], rtol=1e-14)
    assert rep["min_closure"] == pytest.approx(f_pp, rel=1e-14)
    assert rep["max_closure"] == pytest.approx(1.0, rel=1e-14)
    assert rep["max_recip_rel"] < 1e-12
    # pairs with both products >= 0: the 10 unordered pairs
    assert rep["pairs_compared"] == 10


def test_report_finds_planted_reciprocity_error():
    prims = scenes.parse_xml(scenes.config1_xml()).prims
    vf, _ = _config1_exact()
    vf[2, 0] *= 1.01
    rep = rf.vf_report(vf, rf.areas(prims), min_product=1e-9)
    assert (rep["worst_i"], rep["worst_j"]) == (0, 2)
    assert rep["max_recip_rel"] == pytest.approx(0.01 / 1.01, rel=1e-9)
    assert rep["pairs_compared"] == 4          # pairs with a zero entry are skipped
    assert rep["max_closure"] == pytest.approx(1 + 0.01 * vf[2, 0] / 1.01, rel=1e-12)


def test_report_errors():
    with pytest.raises(ValueError):
        rf.vf_report(np.zeros((3, 2)), np.ones(3))
    with pytest.raises(rf.RayFactorError):
        rf.vf_report(np.zeros((0, 0)), np.ones(0))
