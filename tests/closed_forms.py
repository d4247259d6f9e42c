"""Closed-form view factors of the sample scenes and the seam loss that the
self-intersection cut-off EPS = 1e-3 (RayFactorCL/RayFactorGPU.cl:7) causes
where two surfaces touch.

A hit closer than EPS is discarded (cylinderHit's t1 >= EPS, discHit's
t > EPS, ...).  A ray that leaves surface i close to an edge it shares with
surface j reaches j sooner than EPS and passes through it; in these scenes it
then leaves the enclosure.  To first order in EPS, for a straight edge of
length L where the faces meet at the interior dihedral angle alpha:

    a point at distance x from the edge is x sin(alpha) from j's plane; a ray
    of direction d reaches that plane after x sin(alpha) / (d . v), where v is
    the unit vector at angle alpha from the normal n_i, pointing across j
    (d . v > 0 for the rays that reach j).  It is lost when that is < EPS, so
    averaged over x:

        loss_ij = L EPS K(alpha) / (A_i sin(alpha)),    K(b) = E[max(0, d . v)]

    over cosine-weighted d.  With the clamped-cosine integral
    int max(0, d.n) max(0, d.v) dw = (2/3) ((pi - b) cos b + sin b)  (b = angle
    between n and v) and the pdf cos(theta)/pi,

        K(b) = 2 ((pi - b) cos b + sin b) / (3 pi);    K(pi/2) = 2 / (3 pi).

The remainder is O(EPS^2 / R) (edge curvature, second order): below 1e-5 for
these unit-size scenes.  Self-hits of a concave wall within EPS need
cos(theta) < EPS / 2R, probability (EPS / 2R)^2 = 2.5e-7: also in the remainder.

Variance: the 4 rays of a work item are rotations of one random block
(RayFactorGPU.cl:502-506), so the standard error comes from per-work-item hit
counts (sample variance over the work items), not from a binomial over rays.
"""
from __future__ import annotations

import math

import numpy as np

EPS = 1e-3
REMAINDER = 1e-5      # O(EPS^2) bound of the seam-loss expansion for these scenes


def K(b: float) -> float:
    """E[max(0, d . v)] for a cosine-weighted direction d and a unit v at angle b from the normal."""
    return 2.0 * ((math.pi - b) * math.cos(b) + math.sin(b)) / (3.0 * math.pi)


def seam_loss(edge_len: float, area_i: float, alpha: float, eps: float = EPS) -> float:
    """First-order view-factor loss of emitter i toward a surface meeting it at
    dihedral angle alpha along an edge of length edge_len."""
    return edge_len * eps * K(alpha) / (area_i * math.sin(alpha))


def c122_quadrature() -> float:
    """Howell C-122: sphere (R = 1) with its centre at height 1.5 over the corner
    of a 2 x 2 rectangle; F_sphere->rect = (1/4pi) int h / d^3 dA."""
    from numpy.polynomial.legendre import leggauss
    x, w = leggauss(200)
    X = 1 - (x + 1)          # [-1, 1] (rectangle), corner at (1,1)
    dx = (X[:, None] - 1) ** 2 + (X[None, :] - 1) ** 2
    d = np.sqrt(dx + 1.5 ** 2)
    return float((w[:, None] * w[None, :] * 1.5 / d ** 3).sum() / (4 * math.pi))


PI = math.pi
RT2 = math.sqrt(2.0)
# (scene, emitter i, receiver j, closed form, first-order seam loss of (i, j))
CASES = [
    # Howell C-14, two 2 x 2 squares on a common edge (L = 2, A = 4) at 90 deg
    ("input-C14", 0, 1, 0.2000438, seam_loss(2.0, 4.0, PI / 2)),
    ("input-C14-BV", 0, 1, 0.2000438, seam_loss(2.0, 4.0, PI / 2)),       # the same through runBVGPU
    ("input-C47", 0, 1, 0.3819660, 0.0),                                  # disc -> coaxial annulus
    ("input-C52", 0, 1, 0.1534047, 0.0),                                  # ring -> coaxial ring
    ("input-C52-BV", 0, 1, 0.1534047, 0.0),                               # the same through runBVGPU
    # C-109: unit disc under a cone of height 1 (slant sqrt 2); rim L = 2 pi, 45 deg wedge
    ("input-C109", 1, 0, 1.0, seam_loss(2 * PI, PI, PI / 4)),
    ("input-C109", 0, 1, 1 / RT2, seam_loss(2 * PI, PI * RT2, PI / 4)),
    # config 1: closed cylinder r = 1, H = 2 (wall 4 pi, discs pi), rims of 2 pi at 90 deg
    ("config1", 1, 2, 0.1715729, 0.0),                                    # end disc -> opposite end disc
    ("config1", 1, 0, 0.8284271, seam_loss(2 * PI, PI, PI / 2)),          # end disc -> wall
    ("config1", 0, 1, 0.2071068, seam_loss(2 * PI, 4 * PI, PI / 2)),      # wall -> end disc
    ("config1", 0, 0, 0.5857864, 0.0),                                    # wall -> wall (loss in REMAINDER)
    ("config1", 3, 4, 0.3819660, 0.0),                                    # coaxial unit discs, spacing 1
]


def vf_and_sigma(idx: np.ndarray, j: int) -> tuple[float, float]:
    """View factor toward j and its standard error from a per-ray trace
    (4 consecutive rays per work item)."""
    k = (idx.reshape(-1, 4) == j).sum(axis=1).astype(np.float64)
    n = len(k)
    return float(k.mean() / 4.0), float(k.std(ddof=1) / 4.0 / math.sqrt(n))


def check(vf: float, sigma: float, exact: float, loss: float, nsig: float = 3.0) -> bool:
    """The measured VF equals the closed form less the seam loss within nsig
    standard errors (plus the expansion's remainder)."""
    return abs(vf - (exact - loss)) <= nsig * sigma + REMAINDER
