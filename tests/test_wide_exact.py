"""The GPU exact stage for general doubles, compiled for the host: each
pair's 18 doubles are moved into an integer frame (a common power of two,
and a translation by triangle A's first vertex when the coordinates would
not fit 61 bits), decided and constructed on int128/256-bit numbers, and
converted back. Every pair must either be left to the host residual stage
(only when its frame cannot fit) or agree with the oracle exactly: the same
verdict, coplanar flag and exact segments as reduced rationals.

Cases: the golden fixtures (tests/golden/tritri_cases.json), random pairs
of general doubles near each other, pairs sharing a vertex or an edge,
exactly coplanar pairs on a dyadic plane and the same displaced by one ulp,
pairs far from the origin (the translated frame), and pairs with a
coordinate many binades below the others (the residual path)."""
import json
import os
import random
import subprocess
from fractions import Fraction

import pytest

import oracle_py
from gi_test_util import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "wide_test.cpp")
CSRC = os.path.join(ROOT, "gpu-cubql-intersector-standalone_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
GOLDEN = os.path.join(ROOT, "tests", "golden", "tritri_cases.json")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("wide") / "wide_test")
    b = subprocess.run([HIPCC, "-std=c++17", "-O2", "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        "-I" + CSRC, SRC, "-o", exe], capture_output=True, text=True)
    assert b.returncode == 0, b.stdout + b.stderr
    return exe


def _run(exe, cases):
    text = "\n".join(" ".join(float(x).hex() for x in a + b) for a, b in cases) + "\n"
    r = subprocess.run([exe], input=text, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    out, lines, i = [], r.stdout.splitlines(), 0
    while i < len(lines):
        t = lines[i].split()
        i += 1
        if t[0] == "R":
            out.append(None)
            continue
        assert t[0] == "P", t
        hit, cop, deg, nseg, exp2, shifted, npred = (int(x) for x in t[1:8])
        segs = []
        for _ in range(nseg):
            w = lines[i].split()[1:]
            i += 1
            nums = []
            for e in range(8):
                v = sum(int(w[4 * e + k], 16) << (64 * k) for k in range(4))
                nums.append(v - (1 << 256) if v >> 255 else v)
            s = Fraction(2) ** exp2
            p = tuple(Fraction(nums[k], nums[3]) * s for k in range(3))
            q = tuple(Fraction(nums[4 + k], nums[7]) * s for k in range(3))
            segs.append((p, q) if p <= q else (q, p))
        out.append(dict(hit=bool(hit), coplanar=bool(cop), degenerate=bool(deg), segments=sorted(segs),
                        shifted=bool(shifted), npred=npred))
    assert len(out) == len(cases)
    return out


def _frame_bits(a, b):
    """Bits of the largest coordinate of the pair's frame before and after the
    translation by a's first vertex (exact Python integers)."""
    vals = [Fraction(x) for x in a + b]

    def low_exp(v):  # exponent of the lowest set bit of a dyadic rational
        n = abs(v.numerator)
        return (n & -n).bit_length() - 1 - (v.denominator.bit_length() - 1)

    nz = [v for v in vals if v != 0]
    e = min(low_exp(v) for v in nz) if nz else 0
    X = [int(v / Fraction(2) ** e) for v in vals]
    k0 = max(abs(x).bit_length() for x in X)
    k1 = max(abs(X[i] - X[i % 3]).bit_length() for i in range(18))
    return k0, k1


def _check(exe, cases, min_gpu=0.97):
    got = _run(exe, cases)
    n_gpu = 0
    for (a, b), g in zip(cases, got):
        k0, k1 = _frame_bits(a, b)
        if g is None:
            # left to the host only when the frame cannot hold the pair
            assert k0 > 61, (a, b, k0)
            continue
        n_gpu += 1
        assert g["shifted"] == (k0 > 61), (a, b)
        assert not g["shifted"] or k1 <= 61, (a, b)
        ref = oracle_py.tritri(a, b, use_filter=False)
        assert g["hit"] == ref["intersect"], (a, b)
        if g["hit"]:
            assert g["coplanar"] == ref["coplanar"], (a, b)
            assert g["segments"] == ref["segments"], (a, b)
    assert n_gpu >= min_gpu * len(cases)
    return got


def test_golden_fixtures(driver):
    with open(GOLDEN) as f:
        cases = [([float.fromhex(x) for x in c["a"]], [float.fromhex(x) for x in c["b"]])
                 for c in json.load(f)["cases"]]
    _check(driver, cases)


def _tri_near(rng, c, size):
    return [c[k] + rng.uniform(-size, size) for _ in range(3) for k in range(3)]


def test_random_general_doubles(driver):
    rng = random.Random(21)
    cases = []
    for _ in range(1500):
        c = [rng.uniform(-1, 1) for _ in range(3)]
        cases.append((_tri_near(rng, c, 0.05), _tri_near(rng, c, 0.05)))
    got = _check(driver, cases)
    assert sum(g is not None and g["hit"] for g in got) > 200


def test_shared_vertices_edges_and_coplanar(driver):
    rng = random.Random(22)
    cases = []
    for i in range(600):
        a = _tri_near(rng, [rng.uniform(-1, 1) for _ in range(3)], 0.1)
        b = _tri_near(rng, a[:3], 0.1)
        if i % 3 == 0:
            b[0:3] = a[0:3]          # a shared vertex
        elif i % 3 == 1:
            b[0:6] = a[3:9]          # a shared edge
        cases.append((a, b))
    def plane_pt():
        # on the dyadic plane z = x/2 + y/4 + 1/8: exact in doubles for x, y on a 2^-20 grid
        x, y = rng.randint(-2 ** 20, 2 ** 20) / 2 ** 20, rng.randint(-2 ** 20, 2 ** 20) / 2 ** 20
        return [x, y, x / 2 + y / 4 + 0.125]

    for i in range(400):
        a = plane_pt() + plane_pt() + plane_pt()
        b = plane_pt() + plane_pt() + plane_pt()
        if i % 4 == 1:  # b's first vertex one ulp off the plane
            b[2] = b[2] + (abs(b[2]) or 1.0) * 2.0 ** -52
        elif i % 4 == 2:  # a small b inside a's plane near a's first vertex
            b = sum(([a[0] + dx / 2 ** 12, a[1] + dy / 2 ** 12, 0.0] for dx, dy in ((0, 0), (3, 1), (1, 3))), [])
            for j in range(3):
                b[3 * j + 2] = b[3 * j] / 2 + b[3 * j + 1] / 4 + 0.125
        cases.append((a, b))
    got = _check(driver, cases)
    assert sum(g is not None and g["hit"] and g["coplanar"] for g in got) > 20


def test_far_from_origin_takes_the_translated_frame(driver):
    rng = random.Random(23)
    cases = []
    for _ in range(500):
        # two axes far out (~2^20), one near zero (low bits ~2^-60): the
        # unshifted frame would need ~80 bits, the translated one ~50
        c = [rng.uniform(1e5, 1e6), rng.uniform(1e5, 1e6), rng.uniform(-0.02, 0.02)]
        cases.append((_tri_near(rng, c, 1e-3), _tri_near(rng, c, 1e-3)))
    got = _check(driver, cases)
    assert sum(g is not None and g["shifted"] for g in got) > 400
    assert sum(g is not None and g["hit"] for g in got) > 50


def test_rotated_grid_coordinates(driver):
    # integer-grid vertices (2^24 grid) moved by a rotation without rounding:
    # the unsnapped rigid-transform case of the serving and transform paths
    import math
    rng = random.Random(24)
    th = 0.7
    R = [[math.cos(th), -math.sin(th), 0.0], [math.sin(th), math.cos(th), 0.0], [0.0, 0.0, 1.0]]
    cases = []
    for _ in range(500):
        c = [rng.randint(2 ** 21, 2 ** 23) for _ in range(3)]
        a = [c[k] + rng.randint(-4000, 4000) for _ in range(3) for k in range(3)]
        b = [c[k] + rng.randint(-4000, 4000) for _ in range(3) for k in range(3)]
        b = [((R[k][0] * b[3 * j] + R[k][1] * b[3 * j + 1]) + R[k][2] * b[3 * j + 2]) for j in range(3)
             for k in range(3)]
        cases.append((a, b))
    _check(driver, cases)


def test_tiny_coordinates_leave_the_pair_to_the_host(driver):
    rng = random.Random(25)
    cases = []
    for i in range(200):
        c = [rng.uniform(-1, 1) for _ in range(3)]
        a, b = _tri_near(rng, c, 0.05), _tri_near(rng, c, 0.05)
        (a if i % 2 else b)[rng.randrange(9)] = rng.uniform(1, 2) * 2.0 ** -rng.randint(20, 80)
        cases.append((a, b))
    cases.append(([1e300, 0, 0, 0, 1e300, 0, 0, 0, 1e300], [1e-300, 0, 0, 0, 1e-300, 0, 0, 0, 1e-300]))
    got = _run(driver, cases)
    assert sum(g is None for g in got) > 50   # most such pairs need the host
    _check(driver, cases, min_gpu=0.0)
