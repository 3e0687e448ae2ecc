"""Writes tests/golden/tritri_cases.json: triangle pairs with their exact
answers from the independent LP pin (lp_pin.py).

Inputs:
* the one triangle A and six B triangles the reference keeps as a known-answer
  candidate (apps/archive/pair_check.cpp:11-55); the reference records no
  expected output for them, so the answers here come from the exact pin;
* seeded small-integer pairs on a tiny grid (shared vertices, shared edges,
  coplanar overlaps, collinear and point triangles dominate);
* seeded double pairs, including coplanar configurations displaced by one ulp.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import random
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import lp_pin  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tritri_cases.json")

PAIR_CHECK_A = [
    (0.26442988429999997, 0.56879958470000003, 0.71239627849999998),
    (0.24675713530000001, 0.48177735389999998, 0.75283349430000002),
    (0.35325841520000001, 0.53954155599999998, 0.78478380130000003),
]
PAIR_CHECK_B = [
    [(0.28172562919999999, 0.57147973100000005, 0.77439370819999997),
     (0.32367477150000001, 0.53441684369999998, 0.76945816600000005),
     (0.31350119389999997, 0.55834523759999999, 0.82075131040000004)],
    [(0.32367477150000001, 0.53441684369999998, 0.76945816600000005),
     (0.32506398180000001, 0.57775415460000001, 0.78167728790000002),
     (0.31350119389999997, 0.55834523759999999, 0.82075131040000004)],
    [(0.25024791400000002, 0.60007261690000002, 0.7689953383),
     (0.32367477150000001, 0.53441684369999998, 0.76945816600000005),
     (0.28172562919999999, 0.57147973100000005, 0.77439370819999997)],
    [(0.32367477150000001, 0.53441684369999998, 0.76945816600000005),
     (0.29182013979999999, 0.59704641700000005, 0.74041401610000002),
     (0.32506398180000001, 0.57775415460000001, 0.78167728790000002)],
    [(0.31132459089999998, 0.54386790289999998, 0.77399981799999995),
     (0.29182013979999999, 0.59704641700000005, 0.74041401610000002),
     (0.32367477150000001, 0.53441684369999998, 0.76945816600000005)],
    [(0.25024791400000002, 0.60007261690000002, 0.7689953383),
     (0.31132459089999998, 0.54386790289999998, 0.77399981799999995),
     (0.32367477150000001, 0.53441684369999998, 0.76945816600000005)],
]


def frac_str(f):
    return f"{f.numerator}/{f.denominator}"


def case(name, A, B):
    A = [tuple(float(x) for x in p) for p in A]
    B = [tuple(float(x) for x in p) for p in B]
    hit, segs = lp_pin.tritri(A, B)
    return {
        "name": name,
        "a": [x.hex() for p in A for x in p],
        "b": [x.hex() for p in B for x in p],
        "intersect": hit,
        "coplanar": lp_pin.coplanar(A, B),
        "segments": [[[frac_str(c) for c in P], [frac_str(c) for c in Q]] for P, Q in segs],
    }


def main():
    cases = []
    for i, B in enumerate(PAIR_CHECK_B):
        cases.append(case(f"pair_check_{i + 1}", PAIR_CHECK_A, B))
    hand = {
        "shared_vertex": ([(0, 0, 0), (4, 0, 0), (0, 4, 0)], [(0, 0, 0), (0, 0, 4), (-4, 1, 2)]),
        "shared_edge_fold": ([(0, 0, 0), (4, 0, 0), (0, 4, 0)], [(0, 0, 0), (4, 0, 0), (1, 1, 3)]),
        "coplanar_inside": ([(0, 0, 0), (8, 0, 0), (0, 8, 0)], [(1, 1, 0), (3, 1, 0), (1, 3, 0)]),
        "coplanar_hexagon": ([(0, 0, 0), (9, 0, 0), (0, 9, 0)], [(6, 6, 0), (-3, 6, 0), (6, -3, 0)]),
        "coplanar_touch_edge": ([(0, 0, 0), (4, 0, 0), (0, 4, 0)], [(1, 0, 0), (3, 0, 0), (2, -3, 0)]),
        "coplanar_touch_point": ([(0, 0, 0), (4, 0, 0), (0, 4, 0)], [(4, 0, 0), (6, 1, 0), (5, -2, 0)]),
        "coplanar_disjoint": ([(0, 0, 0), (4, 0, 0), (0, 4, 0)], [(3, 3, 0), (6, 3, 0), (3, 6, 0)]),
        "piercing": ([(0, 0, 0), (4, 0, 0), (0, 4, 0)], [(1, 1, -2), (1, 1, 2), (2, 2, 1)]),
        "edge_in_plane": ([(0, 0, 0), (4, 0, 0), (0, 4, 0)], [(-1, 1, 0), (5, 1, 0), (2, 2, 3)]),
        "vertex_on_face": ([(0, 0, 0), (4, 0, 0), (0, 4, 0)], [(1, 1, 0), (1, 1, 3), (2, 3, 3)]),
        "parallel_planes": ([(0, 0, 0), (4, 0, 0), (0, 4, 0)], [(0, 0, 1), (4, 0, 1), (0, 4, 1)]),
        "point_tri_on_face": ([(0, 0, 0), (4, 0, 0), (0, 4, 0)], [(1, 1, 0), (1, 1, 0), (1, 1, 0)]),
        "point_tri_off": ([(0, 0, 0), (4, 0, 0), (0, 4, 0)], [(1, 1, 1), (1, 1, 1), (1, 1, 1)]),
        "segment_tri_crossing": ([(0, 0, 0), (4, 0, 0), (0, 4, 0)], [(1, 1, -1), (1, 1, 1), (1, 1, 0)]),
        "segment_tri_in_plane": ([(0, 0, 0), (4, 0, 0), (0, 4, 0)], [(-1, 1, 0), (5, 1, 0), (2, 1, 0)]),
        "segment_segment_cross": ([(0, 0, 0), (4, 4, 0), (2, 2, 0)], [(0, 4, 0), (4, 0, 0), (1, 3, 0)]),
        "segment_segment_overlap": ([(0, 0, 0), (4, 0, 0), (2, 0, 0)], [(3, 0, 0), (6, 0, 0), (5, 0, 0)]),
        "segment_segment_skew": ([(0, 0, 0), (4, 0, 0), (2, 0, 0)], [(2, -1, 1), (2, 1, 1), (2, 0, 1)]),
        "point_point_equal": ([(1, 2, 3)] * 3, [(1, 2, 3)] * 3),
        "point_point_apart": ([(1, 2, 3)] * 3, [(1, 2, 4)] * 3),
    }
    for name, (A, B) in hand.items():
        cases.append(case(name, A, B))
    rng = random.Random(20260815)
    for i in range(300):
        R = rng.choice([1, 2, 3])
        A = [[rng.randint(-R, R) for _ in range(3)] for _ in range(3)]
        B = [[rng.randint(-R, R) for _ in range(3)] for _ in range(3)]
        if i % 3 == 0:
            for p in A + B:
                p[2] = 0
        cases.append(case(f"grid_{i}", A, B))
    for i in range(60):
        A = [[rng.uniform(-1, 1) for _ in range(3)] for _ in range(3)]
        B = [[rng.uniform(-1, 1) for _ in range(3)] for _ in range(3)]
        cases.append(case(f"double_{i}", A, B))
    for i in range(40):
        # coplanar integer configuration scaled by a non-dyadic factor and
        # one coordinate moved by one ulp: filters must defer, exact decides
        s = rng.choice([0.1, 0.3, 1.7, 3.3e-5])
        A = [[rng.randint(-3, 3) * s for _ in range(2)] + [0.25] for _ in range(3)]
        B = [[rng.randint(-3, 3) * s for _ in range(2)] + [0.25] for _ in range(3)]
        k = rng.randrange(3)
        B[k][2] = float.fromhex((0.25).hex()) + rng.choice([-1, 1]) * 2.0 ** -54
        cases.append(case(f"near_coplanar_{i}", A, B))
    with open(OUT, "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "cases": cases}, f, indent=0)
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
