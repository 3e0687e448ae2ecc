"""The GPU's Morton order and BVH, read back through gi_get_tree, against
Karras' definition (Karras 2012, "Maximizing Parallelism in the
Construction of BVHs, Octrees, and k-d Trees") restated below with numpy
over the GPU's own sorted keys: the keys must be ordered and the ids a
permutation, and every internal node's left-child word (with its right-leaf
bit), rope and leaf range, and every leaf's rope, must equal those of a
plain Karras build (ties between equal keys broken by the index). The kernels
that build it: the staging launch's bucket counts, the bucket sort
(k_msd_scatter, k_msd_buckets) or, for small meshes, the one-workgroup radix
sort; k_build_chunks (in-chunk nodes by binary lifting over adjacent common
prefixes) and k_top_boxes (nodes whose range leaves a chunk, by 16-way
searches). The reference's BVH (cuBQL's level-cut builder,
include/third-party/cubql/level_cut_builder.h:16-364) is internal state whose
shape never reaches the output, so the tree is held to Karras' definition
rather than to it; the candidate sets, which do not depend on the tree, are
checked against the oracle elsewhere."""
import numpy as np
import pytest

from gi_test_util import gi, mg

pytestmark = pytest.mark.gpu

SENTINEL = 0xFFFFFFFF
RIGHT_LEAF = 0x80000000
QUERY_AXIS_BITS = 5  # B's Morton keys (queries only): 5 bits per axis


def _axis_bits(nt):  # a tree's Morton keys: 8 bits per axis up to 2^21 triangles, else 10
    return 8 if nt <= (1 << 21) else 10


def _bit_length(x):
    _, e = np.frexp(x.astype(np.float64))  # exact for integers below 2^53
    return np.where(x > 0, e, 0).astype(np.int64)


def _delta(keys, i, j):
    """Common-prefix length of the composite keys (key, index) of leaves i and
    j (32-bit keys: 32 + the index prefix for equal keys), -1 for j outside."""
    n = len(keys)
    ok = (j >= 0) & (j < n)
    jj = np.clip(j, 0, n - 1)
    x = keys[i].astype(np.int64) ^ keys[jj].astype(np.int64)
    d = np.where(x != 0, 32 - _bit_length(x), 64 - _bit_length(i ^ jj))
    return np.where(ok, d, -1)


def _chain_top(keys, m):
    """The node after every subtree whose last leaf is m - 1: internal node m
    when it extends to the right, else leaf m (node n - 1 + m)."""
    n = len(keys)
    up = (m < n - 1) & (_delta(keys, m, m + 1) > _delta(keys, m, m - 1))
    return np.where(up, m, n - 1 + m)


def karras(keys):
    """Left-child words, ropes, ranges and leaf ropes of the Karras tree over
    sorted keys, every node at once (the per-node searches as in Karras'
    paper: exponential then binary search for the range end, binary search
    for the split)."""
    keys = np.asarray(keys, dtype=np.uint32)
    n = len(keys)
    i = np.arange(n - 1, dtype=np.int64)
    d = np.where(_delta(keys, i, i + 1) > _delta(keys, i, i - 1), 1, -1)
    dmin = _delta(keys, i, i - d)
    lmax = np.full(n - 1, 2, dtype=np.int64)
    while True:
        grow = _delta(keys, i, i + lmax * d) > dmin
        if not grow.any():
            break
        lmax = np.where(grow, 2 * lmax, lmax)
    length = np.zeros(n - 1, dtype=np.int64)
    t = lmax // 2
    while (t >= 1).any():
        step = (t >= 1) & (_delta(keys, i, i + (length + t) * d) > dmin)
        length = np.where(step, length + t, length)
        t = t // 2
    j = i + length * d
    dn = _delta(keys, i, j)
    s = np.zeros(n - 1, dtype=np.int64)
    t = length.copy()
    active = np.ones(n - 1, dtype=bool)
    while active.any():
        t = np.where(active, (t + 1) // 2, t)
        step = active & (_delta(keys, i, i + (s + t) * d) > dn)
        s = np.where(step, s + t, s)
        active &= t > 1
    gamma = i + s * d + np.where(d < 0, -1, 0)
    lo, hi = np.minimum(i, j), np.maximum(i, j)
    left = np.where(lo == gamma, n - 1 + gamma, gamma)
    right = np.where(hi == gamma + 1, n - 1 + gamma + 1, gamma + 1)
    word = left | np.where(right >= n - 1, RIGHT_LEAF, 0)
    rope = np.where(hi == n - 1, SENTINEL, _chain_top(keys, np.minimum(hi + 1, n - 1)))
    k = np.arange(n, dtype=np.int64)
    leaf_rope = np.where(k == n - 1, SENTINEL, _chain_top(keys, np.minimum(k + 1, n - 1)))
    return {"left": word.astype(np.uint32), "rope": rope.astype(np.uint32),
            "range": np.stack([lo, hi], axis=1).astype(np.uint32), "leaf_rope": leaf_rope.astype(np.uint32)}


def _check_order(t, nt):
    keys, ids = t["keys"], t["ids"]
    assert np.all(keys[:-1] <= keys[1:]), "keys out of order"
    assert np.array_equal(np.sort(ids), np.arange(nt, dtype=np.uint32)), "ids are not a permutation"


def _compact(v):
    v = v & 0x09249249
    v = (v | (v >> 2)) & 0x030C30C3
    v = (v | (v >> 4)) & 0x0300F00F
    v = (v | (v >> 8)) & 0x030000FF
    return (v | (v >> 16)) & 0x3FF


def _check_keys(x, t, keys, ids, axis):
    """Every sorted key belongs to its triangle: the Morton cell it names
    is, within one cell per axis (the GPU's keys come from FP32 boxes), that
    of the triangle's box centre in the mesh's bounds."""
    k = keys.astype(np.int64)
    q = np.stack([_compact(k >> 2), _compact(k >> 1), _compact(k)], axis=1)
    p = x[t[ids.astype(np.int64)]]
    c = 0.5 * p.min(axis=1) + 0.5 * p.max(axis=1)
    lo, ext = x.min(axis=0), x.max(axis=0) - x.min(axis=0)
    cells = 1 << axis
    want = np.clip(np.floor((c - lo) / np.where(ext > 0, ext, 1.0) * cells), 0, cells - 1).astype(np.int64)
    bad = np.nonzero(np.abs(q - want).max(axis=1) > 1)[0]
    assert len(bad) == 0, f"{len(bad)} keys do not belong to their triangles, first at {bad[:5]}"


def _check(xa, ta, xb, tb, chunk=None, transform=None):
    eng = gi.Intersector(device=0)
    try:
        if chunk:
            eng.set_tuning(build_chunk=chunk)
        eng.set_mesh(0, xa, ta)
        eng.set_mesh(1, xb, tb)
        if transform:
            eng.set_transform(0, *transform, snap=False)
        eng.build()
        t = eng.get_tree(0, len(ta))
        _check_order(t, len(ta))
        ref = karras(t["keys"])
        for name in ("left", "rope", "range", "leaf_rope"):
            bad = np.nonzero(np.any((t[name] != ref[name]).reshape(len(ref[name]), -1), axis=1))[0]
            assert len(bad) == 0, f"{name} differs at {len(bad)} nodes, first {bad[:5]}"
        tb_ = eng.get_tree(1, len(tb), tree=False)  # B: its query order only
        _check_order(tb_, len(tb))
        if not transform:
            _check_keys(xa, ta, t["keys"], t["ids"], _axis_bits(len(ta)))
        _check_keys(xb, tb, tb_["keys"], tb_["ids"], QUERY_AXIS_BITS)
        return t
    finally:
        eng.close()


@pytest.mark.parametrize("chunk", [256, 512])
def test_bench_meshes_tree_is_karras(chunk):
    # 2^20 triangles a side: bucket sorts of 24-bit (A) and 15-bit (B) keys,
    # including the sphere's pole cells of ~15K keys
    (xa, ta), (xb, tb) = mg.sphere_torus(1024, 512)
    t = _check(xa, ta, xb, tb, chunk=chunk)
    assert len(np.unique(t["keys"])) < len(ta)  # equal keys, ordered by the index tie-break


def test_icosphere_tree_is_karras():
    (xa, ta), (xb, tb) = mg.icosphere_pair(5)
    _check(xa, ta, xb, tb)


def test_moved_icosphere_tree_is_karras():
    # unsnapped rigid motion: keys from the moved doubles
    (xa, ta), (xb, tb) = mg.icosphere_pair(5)
    _check(xa, ta, xb, tb, transform=((17.0, -31.0, 5.0), (0.125, -0.25, 0.5), (0.0, 0.0, 0.0)))


def test_height_field_tree_is_karras():
    (pa, t), (pb, _), _ = mg.height_field_pairs(128)
    _check(pa, t, pb, t)


def test_small_torus_tree_is_karras():
    # under 4,097 triangles: the one-workgroup radix sort and a single chunk
    xs, ts = mg.torus(37, 23)
    (xb, tb), _ = mg.icosphere_pair(3)
    _check(mg.snap(xs), ts, xb, tb)


# Buckets (the top 4 key bits per axis) of more than 4,096 keys are split
# over several workgroups, up to 32 such buckets per sort, through a global
# low-part histogram that each sort leaves cleared.

def test_query_side_split_buckets():
    # the sphere as B: 24 split buckets of 15-bit keys (3 low bits)
    (xs, ts), (xt, tt) = mg.sphere_torus(1024, 512)
    _check(xt, tt, xs, ts)


def test_more_split_buckets_than_slots():
    # 2^21 triangles: 66 buckets over 4,096 keys, the 34 past the slots sorted
    # by one workgroup each
    xs, ts = mg.uv_sphere(2048, 512, 3)
    (xb, tb), _ = mg.icosphere_pair(3)
    _check(mg.snap(xs), ts, xb, tb)


def test_split_buckets_repeated_sorts():
    # one context sorting meshes with and without split buckets in turn (by
    # gi_build after new meshes, then by a query that rebuilds): the slots'
    # histograms and cursors must start every sort at zero
    (xs, ts), (xt, tt) = mg.sphere_torus(1024, 512)
    xs2, ts2 = mg.uv_sphere(2048, 512, 3)
    xs2 = mg.snap(xs2)
    eng = gi.Intersector(device=0)

    def check(xa, ta, xb, tb):
        t = eng.get_tree(0, len(ta), tree=False)
        _check_order(t, len(ta))
        _check_keys(xa, ta, t["keys"], t["ids"], _axis_bits(len(ta)))
        u = eng.get_tree(1, len(tb), tree=False)
        _check_order(u, len(tb))
        _check_keys(xb, tb, u["keys"], u["ids"], QUERY_AXIS_BITS)

    try:
        for xa, ta, xb, tb in ((xs, ts, xt, tt), (xt, tt, xs, ts), (xs2, ts2, xt, tt), (xs, ts, xt, tt)):
            eng.set_mesh(0, xa, ta)
            eng.set_mesh(1, xb, tb)
            eng.build()
            check(xa, ta, xb, tb)
            res = eng.intersect_raw(output_on_device=True, want_segments=False, rebuild=True)
            gi.lib().gi_result_free(gi.C.byref(res))
            check(xa, ta, xb, tb)
    finally:
        eng.close()
