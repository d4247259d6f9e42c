"""Checks of the oracle's nearest-hit and bounding-surface functions
(findObject RayFactorCL/RayFactorGPU.cl:445-472, runBVGPU :578-624) that do
not restate them: constructed rays whose answer is fixed by geometry, and
brute force over the per-primitive hit test.  Each takes the oracle module
``O`` so test_oracle_mutants.py can run it against oracles with planted
mistakes (every pin must fail on at least one of them)."""
from __future__ import annotations

import numpy as np

DISC = 5
INF = float("inf")
I12 = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], dtype=np.float32)


def _prims(n):
    from oracle import PRIM_DTYPE
    p = np.zeros(n, dtype=PRIM_DTYPE)
    p["type"] = DISC
    p["m"] = I12
    p["invm"] = I12
    return p


def _bvs(n):
    from oracle import BV_DTYPE
    b = np.zeros(n, dtype=BV_DTYPE)
    b["type"] = DISC
    b["invm"] = I12
    return b


def _at(p, z, x=0.0):
    """Unit discs at heights z (object -> world: translation)."""
    p["invm"][:, 11] = -np.asarray(z, dtype=np.float32)
    p["m"][:, 11] = np.asarray(z, dtype=np.float32)
    p["invm"][:, 3] = -np.asarray(x, dtype=np.float32)
    p["m"][:, 3] = np.asarray(x, dtype=np.float32)


def _fns(O, mode):
    if mode == "compat":
        return O.find_f32, O.find_bv_f32, O.hit_f32, np.float32
    return O.find_f64, O.find_bv_f64, O.hit_f64, np.float64


def pin_find_tie(O, mode):
    """Strict '<': the lowest index wins an exact tie (Q10); on random rays the
    scan returns the argmin of the per-primitive hit distances."""
    find, _, hit, ft = _fns(O, mode)
    p = _prims(3)
    r = np.array([0.2, 0.1, 3, 0, 0, -1], dtype=ft)
    assert find(p, 0, 3, r) == 0          # three coincident discs: exact tie
    assert find(p, 1, 3, r) == 1
    assert find(p, 2, 2, r) == -1         # empty range
    _at(p, [0.0, 1.0, 2.0])
    rng = np.random.default_rng(3)
    for _ in range(200):
        s = np.concatenate([rng.uniform(-1.2, 1.2, 2), [rng.uniform(-1, 3)]])
        c = rng.normal(size=3)
        c /= np.linalg.norm(c)
        rr = np.concatenate([s, c]).astype(ft)
        ts = [hit(DISC, p["invm"][i], 0, 0, rr) for i in range(3)]
        best = int(np.argmin(ts)) if min(ts) < INF else -1
        assert find(p, 0, 3, rr) == best


def pin_bv_stale_index(O, mode):
    """Q12: an intermediate-check surface sets objectIndex; a later plain
    surface that is nearer wins on the surfaces' t and keeps that stale index
    (RayFactorGPU.cl:596-619).  Control: without the intermediate check the
    winner's child scan gives its own child."""
    _, find_bv, _, ft = _fns(O, mode)
    p = _prims(2)
    _at(p, [1.0, 2.0])
    b = _bvs(2)
    b["invm"] = p["invm"]
    b["settings"] = [1, 0]
    b["objStartIndex"] = [0, 1]
    b["objEndIndex"] = [1, 2]
    r = np.array([0, 0, 3, 0, 0, -1], dtype=ft)
    assert find_bv(p, b, r) == 0
    b["settings"] = [0, 0]
    assert find_bv(p, b, r) == 1


def pin_bv_intermediate_miss(O, mode):
    """An intermediate-check surface whose children the ray misses is not
    taken, however near (RayFactorGPU.cl:598-607): the farther plain surface
    wins and its child is hit.  A ray that misses every surface hits nothing,
    even a primitive outside every surface."""
    _, find_bv, _, ft = _fns(O, mode)
    p = _prims(3)
    _at(p, [1.5, 1.0, 0.5], x=[5.0, 0.0, 0.0])   # child of BV0 off to the side; prim 2 in no BV
    b = _bvs(2)
    b["invm"][:, 11] = [-2.0, -1.0]
    b["settings"] = [1, 0]
    b["objStartIndex"] = [0, 1]
    b["objEndIndex"] = [1, 2]
    r = np.array([0.5, 0, 3, 0, 0, -1], dtype=ft)
    assert find_bv(p, b, r) == 1
    r_out = np.array([3.0, 0, 3, 0, 0, -1], dtype=ft)  # misses both surfaces; would hit nothing anyway
    assert find_bv(p, b, r_out) == -1
    b["invm"][:, 3] = -10.0                          # surfaces moved away: nothing is reachable
    assert find_bv(p, b, r) == -1


def pin_bv_closed_form(O, name, sc, mode, n_wi=1 << 18):
    """A bounding-surface scene reproduces its Howell closed form less the seam
    loss (closed_forms.py), through the runBVGPU semantics."""
    from closed_forms import CASES, check, vf_and_sigma
    _, i, j, exact, loss = next(c for c in CASES if c[0] == name)
    assert sc.useBoundingVolumes
    idx = O.trace(sc.prims, sc.bvs, i, 0, n_wi, mode=mode, use_bvs=True)
    vf, se = vf_and_sigma(idx, j)
    assert check(vf, se, exact, loss), (name, mode, vf, exact - loss, se)
