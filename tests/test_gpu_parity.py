"""GPU parity: the HIP path (through the C ABI) against the oracle.

Bar: identical sorted pair sets, identical candidate counts (exact closed
box overlaps, the set CGAL's box_intersection_d reports), identical exact
segment endpoints (compared as reduced rationals)."""
import numpy as np
import pytest

from gi_test_util import gi, mg, oracle_py, run_gpu

pytestmark = pytest.mark.gpu


def _check(xa, ta, xb, tb, expect_grid, **kw):
    res = run_gpu(xa, ta, xb, tb, **kw)
    ref = oracle_py.intersect(xa, ta, xb, tb, nthreads=8)
    st = res["stats"]
    assert st["grid"] == int(expect_grid)
    np.testing.assert_array_equal(res["pairs"], ref["pairs"])
    assert st["candidates"] == ref["stats"]["n_candidates"]
    assert res["segments"] == ref["segments"]
    assert st["coplanar"] == ref["stats"]["n_coplanar"]
    assert st["degenerate"] == ref["stats"]["n_degenerate"]
    assert st["filter_no"] + st["filter_green"] + st["filter_yellow"] == st["candidates"]
    return res, ref


def test_sphere_torus_small_grid():
    (xa, ta), (xb, tb) = mg.sphere_torus(128, 64)
    res, ref = _check(xa, ta, xb, tb, True)
    assert len(res["pairs"]) > 100


def test_result_output_one_workgroup():
    # 409,600 triangles per mesh give between 4,096 and 16,384 intersecting
    # pairs: the small-result output (one launch: a workgroup per range of
    # keys counts the records below its range, sorts its own in LDS and
    # places them) sorts them and derives their segment offsets
    (xa, ta), (xb, tb) = mg.sphere_torus(640, 320)
    res, ref = _check(xa, ta, xb, tb, True)
    assert 4096 < len(res["pairs"]) <= 16384


def test_result_output_skewed_to_one_key_range():
    # every pair on face 0 of A (a large triangle in the plane z = 0; A's
    # other 1,023 faces are small and far away) against a wall of 2 x 2,500
    # triangles in the plane x = 1 crossing it: 5,000 pairs whose keys all
    # share their top bits, so the small-result output sorts them all in one
    # of its key ranges
    n = 2500
    big = np.array([[-8 * n, -8 * n, 0], [8 * n, -8 * n, 0], [0, 8 * n, 0]], dtype=np.float64)
    k = np.arange(1023, dtype=np.float64)
    far = np.stack([np.stack([1000 + 4 * k, np.zeros_like(k), np.full_like(k, 1000)], 1),
                    np.stack([1002 + 4 * k, np.zeros_like(k), np.full_like(k, 1000)], 1),
                    np.stack([1000 + 4 * k, np.full_like(k, 2), np.full_like(k, 1000)], 1)], 1).reshape(-1, 3)
    xa = np.concatenate([big, far])
    ta = np.arange(len(xa), dtype=np.uint32).reshape(-1, 3)
    y = np.arange(n + 1, dtype=np.float64)
    xb = np.concatenate([np.stack([np.ones_like(y), y, -np.ones_like(y)], 1),
                         np.stack([np.ones_like(y), y, np.ones_like(y)], 1)])
    j = np.arange(n, dtype=np.uint32)
    lo0, lo1, hi0, hi1 = j, j + 1, j + n + 1, j + n + 2
    tb = np.concatenate([np.stack([lo0, lo1, hi1], 1), np.stack([lo0, hi1, hi0], 1)]).astype(np.uint32)
    res, ref = _check(xa, ta, xb, tb, True)
    assert len(res["pairs"]) == 2 * n and np.all(res["pairs"][:, 0] == 0)


def test_result_sort_multi_tile():
    # a 96 x 96 height field against its 90-degree turn crosses along many
    # curves: more than 16,384 pairs, so the result sort runs the multi-tile
    # radix passes, the gather and the multi-block scan
    (p, t), _, (r, _) = mg.height_field_pairs(cells=96)
    res, ref = _check(p, t, r, t, True)
    assert len(res["pairs"]) > 16384


def test_sphere_torus_multi_block_sort_and_chunks():
    # 131,072 triangles per mesh: the Morton sort takes the multi-block radix
    # path (more than 16,384 keys) and the BVH spans 256 build chunks
    (xa, ta), (xb, tb) = mg.sphere_torus(256, 256)
    res, ref = _check(xa, ta, xb, tb, True)
    assert len(res["pairs"]) > 1000


def test_sphere_above_2_21_triangles_with_small_torus():
    # a 2,359,296-triangle sphere: past 2^21 triangles its Morton keys take 10
    # bits per axis (30-bit keys, 4 radix passes) while the 65,536-triangle
    # torus keeps 8 (3 passes), so the two sorts sharing each pass's launches
    # differ in tiles and in passes (the torus drops out of the last one)
    (xa, ta), _ = mg.sphere_torus(1536, 768)
    _, (xb, tb) = mg.sphere_torus(256, 128)
    assert len(ta) > (1 << 21) and len(tb) < (1 << 21)
    res, ref = _check(xa, ta, xb, tb, True)
    assert len(res["pairs"]) > 1000


def test_bench_workload_full_size():
    # the measured workload itself: displaced sphere x rotated torus, 1,048,576
    # triangles each, every pair, candidate count and segment against the oracle
    (xa, ta), (xb, tb) = mg.sphere_torus(1024, 512)
    res, ref = _check(xa, ta, xb, tb, True)
    assert len(res["pairs"]) > 10000


@pytest.mark.parametrize("level", [4, 5])
def test_icosphere_pair_doubles_gpu_exact(level):
    # level 5 is the reference's CPU-runnable case: 20,480 triangles each, raw
    # doubles (off the integer grid): the exact stage runs on the GPU in each
    # pair's integer frame; only pairs whose frame exceeds 61 bits (a
    # coordinate many binades below the pair's others) go to the host
    (xa, ta), (xb, tb) = mg.icosphere_pair(level=level)
    res, ref = _check(xa, ta, xb, tb, False)
    st = res["stats"]
    assert len(res["pairs"]) > 50
    assert st["exact_pairs"] > 0 and st["host_pairs"] <= 0.02 * (st["exact_pairs"] + st["host_pairs"])
    assert st["exact_predicates"] > 0 and st["filter_predicates"] > 0


def test_residual_pairs_merge_with_gpu_pairs():
    # vertices near the plane y = 0, which the intersection curve of the two
    # spheres crosses, moved onto it up to tiny y coordinates (2^-70 .. 2^-90):
    # their pairs' frames exceed 61 bits, so those pairs are decided on the
    # host and merged, in (a, b) order, with the pairs the GPU decided
    (xa, ta), (xb, tb) = mg.icosphere_pair(level=4)
    xa = xa.copy()
    near = np.nonzero(np.abs(xa[:, 1]) < 0.02)[0]
    assert len(near) > 10
    xa[near, 1] = np.ldexp(np.sign(xa[near, 1]) + (xa[near, 1] == 0), -70 - (near % 21))
    res, ref = _check(xa, ta, xb, tb, False)
    st = res["stats"]
    assert st["host_pairs"] > 0 and st["exact_pairs"] > st["host_pairs"]
    assert len(res["pairs"]) > 50


def test_oneshot_entry_point():
    """gi_intersect_oneshot (the reference's runGridForestIntersectionPipeline:
    copy, build and intersect in one call) on grid meshes and on doubles."""
    for (xa, ta), (xb, tb), grid in ((*mg.sphere_torus(96, 48), 1), (*mg.icosphere_pair(level=3), 0)):
        res = gi.intersect_oneshot(xa, ta, xb, tb)
        ref = oracle_py.intersect(xa, ta, xb, tb, nthreads=8)
        assert res["stats"]["grid"] == grid
        np.testing.assert_array_equal(res["pairs"], ref["pairs"])
        assert res["segments"] == ref["segments"]
        assert res["stats"]["candidates"] == ref["stats"]["n_candidates"]
        assert len(res["pairs"]) > 10


def test_height_field_shift_and_rotation():
    (p, t), (s, _), (r, _) = mg.height_field_pairs(cells=48)
    res1, _ = _check(p, t, s, t, True)
    res2, _ = _check(p, t, r, t, True)
    # the 90-degree copy shares every grid vertex: coplanar and degenerate
    # contacts dominate and must be exercised
    assert res2["stats"]["coplanar"] > 0
    assert res2["stats"]["exact_yellow_yes"] > 0


@pytest.mark.parametrize("copy", ["shift", "rot90"])
def test_full_size_height_field_on_sampled_triangles(copy):
    """The degeneracy-heavy pair at its full size (512 x 512 cells, 524,288
    triangles a side; the oracle takes over a minute on the whole pair): the
    GPU result of the whole pair, restricted to every 17th triangle of A,
    equals the oracle's result for those triangles, as a mesh of their own,
    against all of B. A pair's verdict and segment depend on its two
    triangles only, so the restriction is exact."""
    (p, t), (s, _), (r, _) = mg.height_field_pairs(cells=512)
    q = s if copy == "shift" else r
    eng = gi.Intersector(device=0)
    try:
        eng.set_mesh(0, p, t)
        eng.set_mesh(1, q, t)
        res = eng.intersect(decode=False)
    finally:
        eng.close()
    pairs, offs = res["pairs"], res["seg_offsets"].astype(np.int64)
    assert res["stats"]["grid"] == 1
    assert np.all(np.diff(pairs[:, 0].astype(np.int64) * (1 << 32) + pairs[:, 1]) > 0)  # sorted, no repeats
    stride = 17
    sel = np.nonzero(pairs[:, 0] % stride == 0)[0]
    ref = oracle_py.intersect(p, np.ascontiguousarray(t[::stride]), q, t, nthreads=8)
    got = pairs[sel].copy()
    got[:, 0] //= stride
    np.testing.assert_array_equal(got, ref["pairs"])
    assert ref["stats"]["n_coplanar"] > 0  # the sample reaches the coplanar (deferred) construction
    counts = offs[sel + 1] - offs[sel]
    sub_offs = np.concatenate([[0], np.cumsum(counts)])
    idx = np.repeat(offs[sel] - sub_offs[:-1], counts) + np.arange(int(sub_offs[-1]))
    exp2 = res["seg_exp2"][idx] if res["seg_exp2"] is not None else None
    assert gi.decode_segments(res["seg_exact"][idx], exp2, sub_offs) == ref["segments"]


def test_transform_refit_matches_oracle_on_moved_vertices():
    (xa, ta), (xb, tb) = mg.sphere_torus(96, 48)
    eng = gi.Intersector()
    try:
        eng.set_mesh(0, xa, ta)
        eng.set_mesh(1, xb, tb)
        first = eng.intersect()
        c = xb.mean(axis=0)
        eng.set_transform(1, [10.0, -20.0, 35.0], [2.0e5, -1.0e5, 3.0e4], c.tolist(), snap=True)
        moved = eng.intersect()
    finally:
        eng.close()
    # the same motion on the host: CGAL-order affine map, then rounding
    R = mg.euler_matrix(*(np.radians([10.0, -20.0, 35.0])))
    # rebuild the matrix exactly as the library does
    xb2 = gi_apply_rigid(xb, [10.0, -20.0, 35.0], [2.0e5, -1.0e5, 3.0e4], c)
    ref = oracle_py.intersect(xa, ta, np.rint(xb2), tb, nthreads=8)
    np.testing.assert_array_equal(moved["pairs"], ref["pairs"])
    assert moved["segments"] == ref["segments"]
    assert len(first["pairs"]) > 0 and R.shape == (3, 3)


@pytest.mark.parametrize("case", ["sphere_torus", "icospheres"])
def test_transform_unrounded_matches_oracle_bit_for_bit(case):
    """The reference's own transform flow: the moved vertices are not rounded
    (setTransformBoth never rounds). The doubles the library moves must be
    bit-identical to the CGAL-order map evaluated on the host (numpy rounds
    each product and sum separately, in that order), and the pairs and exact
    segments equal the oracle's on those doubles."""
    if case == "sphere_torus":
        (xa, ta), (xb, tb) = mg.sphere_torus(96, 48)
        rot, tr = [10.0, -20.0, 35.0], [2.0e5, -1.0e5, 3.0e4]
    else:
        (xa, ta), (xb, tb) = mg.icosphere_pair(level=4)
        rot, tr = [17.0, 41.0, -23.0], [-0.05, 0.02, 0.01]
    c = xb.mean(axis=0)
    eng = gi.Intersector()
    try:
        eng.set_mesh(0, xa, ta)
        eng.set_mesh(1, xb, tb)
        eng.set_transform(1, rot, tr, c.tolist(), snap=False)
        moved_dev = eng.get_vertices(1, xb.shape[0])
        moved = eng.intersect()
    finally:
        eng.close()
    xb2 = gi_apply_rigid(xb, rot, tr, c)
    assert np.array_equal(moved_dev.view(np.uint64), xb2.view(np.uint64))  # bit for bit
    assert not np.array_equal(xb2, np.rint(xb2))
    ref = oracle_py.intersect(xa, ta, xb2, tb, nthreads=8)
    st = moved["stats"]
    assert st["grid"] == 0
    np.testing.assert_array_equal(moved["pairs"], ref["pairs"])
    assert moved["segments"] == ref["segments"]
    assert st["candidates"] == ref["stats"]["n_candidates"]
    assert st["coplanar"] == ref["stats"]["n_coplanar"] and st["degenerate"] == ref["stats"]["n_degenerate"]
    assert len(moved["pairs"]) > 10
    assert st["host_pairs"] <= 0.02 * (st["exact_pairs"] + st["host_pairs"])


def gi_apply_rigid(x, rot, trans, c):
    import math
    d2r = 3.14159265358979323846 / 180.0
    rx, ry, rz = rot[0] * d2r, rot[1] * d2r, rot[2] * d2r
    cx, sx, cy, sy, cz, sz = math.cos(rx), math.sin(rx), math.cos(ry), math.sin(ry), math.cos(rz), math.sin(rz)
    R = [[cy * cz, sx * sy * cz - cx * sz, cx * sy * cz + sx * sz],
         [cy * sz, sx * sy * sz + cx * cz, cx * sy * sz - sx * cz],
         [-sy, sx * cy, cx * cy]]
    T = [c[r] + trans[r] - (R[r][0] * c[0] + R[r][1] * c[1] + R[r][2] * c[2]) for r in range(3)]
    out = np.empty_like(x)
    for r in range(3):
        out[:, r] = ((R[r][0] * x[:, 0] + R[r][1] * x[:, 1]) + R[r][2] * x[:, 2]) + T[r]
    return out


def test_empty_and_single_triangle_meshes():
    xa = np.array([[0, 0, 0], [4, 0, 0], [0, 4, 0]], dtype=np.float64)
    ta = np.array([[0, 1, 2]], dtype=np.uint32)
    xb = np.array([[1, 1, -2], [1, 1, 2], [2, 2, 1]], dtype=np.float64)
    res, ref = _check(xa, ta, xb, ta, True)
    assert len(res["pairs"]) == 1
    empty = run_gpu(xa, np.zeros((0, 3), dtype=np.uint32), xb, ta)
    assert len(empty["pairs"]) == 0


@pytest.mark.parametrize("na", [2, 3, 5])
def test_tiny_tree_against_many_queries(na):
    # a tree of 2-5 triangles (its root's children are leaves or one level
    # above them) walked by full waves of queries: the wave-shared descent
    # stops at once or after one level, and the lanes finish the walk
    (xa, ta), (xb, tb) = mg.sphere_torus(96, 48)
    hits = oracle_py.intersect(xa, ta, xb, tb, nthreads=8, want_segments=False)["pairs"]
    picked = np.unique(hits[:, 0])[:: max(1, len(np.unique(hits[:, 0])) // na)][:na]  # A triangles that cross B
    sel = np.ascontiguousarray(ta[picked])
    used, inv = np.unique(sel, return_inverse=True)
    xa2, ta2 = np.ascontiguousarray(xa[used]), inv.reshape(-1, 3).astype(np.uint32)
    res, _ = _check(xa2, ta2, xb, tb, True)
    assert len(res["pairs"]) >= na


def test_queries_that_miss_the_tree():
    # B far beside A: every wave's union box misses A's root (the descent ends
    # the whole wave); then B inside A's box but off its surface: unions that
    # miss both children of a node end there too
    (xa, ta), (xb, tb) = mg.sphere_torus(96, 48)
    far = np.rint(0.3 * (xb - xb.mean(axis=0)) + 0.4 * mg.GRID_SCALE)  # a small torus in the grid's corner
    assert far.min() >= 0 and far.max() < xa.min()
    res, ref = _check(xa, ta, far, tb, True)
    assert len(res["pairs"]) == 0 and res["stats"]["candidates"] == 0
    c = xa.mean(axis=0)
    small = np.rint(c + 0.2 * (xb - xb.mean(axis=0)))  # a torus a fifth the size, at the sphere's centre
    res, ref = _check(xa, ta, small, tb, True)
    assert len(res["pairs"]) == 0
