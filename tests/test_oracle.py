"""The oracle pinned against the golden fixtures and independent exact math.

The fixtures (tests/golden/tritri_cases.json) carry answers computed by an
LP/basic-feasible-solution enumeration in exact rationals (lp_pin.py), which
shares no code or algorithm with the oracle."""
import json
import os
import random
from fractions import Fraction

import numpy as np
import pytest

import lp_pin
import oracle_py

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "tritri_cases.json")


def _parse_frac(s):
    n, d = s.split("/")
    return Fraction(int(n), int(d))


def _golden_cases():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("use_filter", [False, True])
def test_oracle_matches_golden(use_filter):
    for c in _golden_cases():
        a = [float.fromhex(x) for x in c["a"]]
        b = [float.fromhex(x) for x in c["b"]]
        r = oracle_py.tritri(a, b, use_filter=use_filter)
        assert r["intersect"] == c["intersect"], c["name"]
        want = sorted(tuple(tuple(_parse_frac(x) for x in P) for P in seg) for seg in c["segments"])
        assert r["segments"] == want, c["name"]
        if c["intersect"]:
            assert r["coplanar"] == c["coplanar"], c["name"]


def test_pair_check_known_answers_are_all_intersecting():
    # apps/archive/pair_check.cpp prints INTERSECT/NO INTERSECT for six B
    # triangles without recording an answer; exact arithmetic says all six meet A.
    cases = [c for c in _golden_cases() if c["name"].startswith("pair_check_")]
    assert len(cases) == 6 and all(c["intersect"] for c in cases)


def test_oracle_vs_lp_pin_fresh_small_grid():
    rng = random.Random(7)
    for i in range(250):
        R = rng.choice([1, 2])
        A = [[rng.randint(-R, R) for _ in range(3)] for _ in range(3)]
        B = [[rng.randint(-R, R) for _ in range(3)] for _ in range(3)]
        if i % 2:
            for p in A + B:
                p[0] = p[1] - p[2]  # all points on one plane, not axis aligned
        hit, segs = lp_pin.tritri(A, B)
        r = oracle_py.tritri(sum(A, []), sum(B, []), use_filter=True)
        assert r["intersect"] == hit and r["segments"] == segs, (A, B)


def test_orient3d_sign_and_filter_soundness():
    rng = random.Random(3)
    for i in range(2000):
        if i % 2:
            pts = [[rng.randint(-5, 5) * 0.125 for _ in range(3)] for _ in range(4)]
        else:
            pts = [[rng.uniform(-1, 1) for _ in range(3)] for _ in range(4)]
            if i % 4 == 0:  # nearly coplanar: d close to the plane of a, b, c
                t, u = rng.random(), rng.random()
                pts[3] = [pts[0][k] + t * (pts[1][k] - pts[0][k]) + u * (pts[2][k] - pts[0][k])
                          for k in range(3)]
        exact = lp_pin.orient3d(*pts)
        s = (exact > 0) - (exact < 0)
        assert oracle_py.orient3d_sign(*pts) == s
        f = oracle_py.orient3d_filter(*pts)
        assert f == 0 or f == s
        # antisymmetry under swapping two points, invariance under cyclic shift of a,b,c
        a, b, c, d = pts
        assert oracle_py.orient3d_sign(b, a, c, d) == -s
        assert oracle_py.orient3d_sign(b, c, a, d) == s


def _random_soup(rng, n, scale, grid=False):
    xyz = rng.uniform(-1, 1, size=(3 * n, 3))
    if grid:
        xyz = np.round(xyz * 16)
    xyz += np.repeat(rng.uniform(-scale, scale, size=(n, 3)), 3, axis=0)
    tri = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    return xyz, tri


@pytest.mark.parametrize("grid", [False, True])
def test_candidates_sweep_equals_brute_force(grid):
    rng = np.random.default_rng(11)
    xA, tA = _random_soup(rng, 700, 8 if not grid else 60, grid)
    xB, tB = _random_soup(rng, 650, 8 if not grid else 60, grid)
    sweep = oracle_py.candidates(xA, tA, xB, tB, method=0, nthreads=4)
    brute = oracle_py.candidates(xA, tA, xB, tB, method=1, nthreads=4)
    assert len(sweep) > 0
    np.testing.assert_array_equal(sweep, brute)
    # numpy brute force over exact double boxes
    def boxes(x, t):
        v = x[t.astype(np.int64)]
        return v.min(axis=1), v.max(axis=1)
    la, ha = boxes(xA, tA)
    lb, hb = boxes(xB, tB)
    ov = np.all((la[:, None, :] <= hb[None, :, :]) & (lb[None, :, :] <= ha[:, None, :]), axis=2)
    ia, ib = np.nonzero(ov)
    np.testing.assert_array_equal(sweep, np.stack([ia, ib], axis=1).astype(np.uint32))


def test_intersect_pipeline_filter_equals_exact_and_lp():
    rng = np.random.default_rng(5)
    xA, tA = _random_soup(rng, 300, 24, grid=True)
    xB, tB = _random_soup(rng, 300, 24, grid=True)
    r1 = oracle_py.intersect(xA, tA, xB, tB, use_filter=True)
    r0 = oracle_py.intersect(xA, tA, xB, tB, use_filter=False)
    np.testing.assert_array_equal(r1["pairs"], r0["pairs"])
    assert r1["segments"] == r0["segments"]
    assert r1["stats"]["n_candidates"] == r0["stats"]["n_candidates"]
    assert len(r1["pairs"]) > 10
    # a sample of reported pairs and of rejected candidates agree with the LP pin
    cands = oracle_py.candidates(xA, tA, xB, tB)
    hits = {tuple(p): i for i, p in enumerate(r1["pairs"].tolist())}
    hit_list = [tuple(c) for c in cands.tolist() if tuple(c) in hits]
    miss_list = [tuple(c) for c in cands.tolist() if tuple(c) not in hits]
    pick = random.Random(1)
    sample = pick.sample(hit_list, min(60, len(hit_list))) + pick.sample(miss_list, min(60, len(miss_list)))
    for a, b in sample:
        A = xA[tA[a].astype(np.int64)].tolist()
        B = xB[tB[b].astype(np.int64)].tolist()
        hit, segs = lp_pin.tritri(A, B)
        assert hit == ((a, b) in hits)
        if hit:
            assert r1["segments"][hits[(a, b)]] == segs
