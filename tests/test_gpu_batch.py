"""Batches (gi_set_batch / gi_intersect_batch): every mesh pair's result must
equal what the single-pair entry point returns for that pair alone, and the
oracle's on the same arrays."""
import numpy as np
import pytest

from gi_test_util import gi, mg, oracle_py

pytestmark = pytest.mark.gpu


def _single(eng, pair):
    (xa, ta), (xb, tb) = pair
    eng.set_mesh(0, xa, ta)
    eng.set_mesh(1, xb, tb)
    return eng.intersect()


def _edge_pairs():
    """empty side, one triangle against a mesh, a mesh against itself
    (coplanar everywhere), a mesh against a disjoint one"""
    (xa, ta), (xb, tb) = mg.batch_workload(1, seed=9, nmin=300, nmax=300)[0]
    empty = (np.zeros((0, 3)), np.zeros((0, 3), dtype=np.uint32))
    one = (xa[ta[:1].ravel()], np.arange(3, dtype=np.uint32).reshape(1, 3))
    far = (xb + 8.0 * (1 << 22), tb)
    return [(empty, (xb, tb)), ((xa, ta), empty), (one, (xa, ta)), ((xa, ta), (xa, ta)), ((xa, ta), far)]


def _check_batch(pairs, n_oracle=None, rebuild_twice=True):
    eng = gi.Intersector()
    try:
        eng.set_batch(pairs)
        got, stats = eng.intersect_batch()
        if rebuild_twice:
            again, _ = eng.intersect_batch(rebuild=True)
            for g, a in zip(got, again):
                np.testing.assert_array_equal(g["pairs"], a["pairs"])
                assert g["segments"] == a["segments"]
        # the same context goes back to single pairs
        singles = [_single(eng, p) for p in pairs]
    finally:
        eng.close()
    assert len(got) == len(pairs)
    for i, (g, s) in enumerate(zip(got, singles)):
        np.testing.assert_array_equal(g["pairs"], s["pairs"], err_msg=f"pair {i}")
        assert g["segments"] == s["segments"], f"pair {i}"
    assert sum(len(g["pairs"]) for g in got) == stats["pairs"]
    checked = 0
    for i, ((xa, ta), (xb, tb)) in enumerate(pairs):
        if (n_oracle is not None and checked >= n_oracle) or len(ta) == 0 or len(tb) == 0:
            continue
        ref = oracle_py.intersect(xa, ta, xb, tb, nthreads=8)
        np.testing.assert_array_equal(got[i]["pairs"], ref["pairs"], err_msg=f"pair {i}")
        assert got[i]["segments"] == ref["segments"], f"pair {i}"
        checked += 1
    return got, stats


def test_grid_batch_matches_single_pairs_and_oracle():
    pairs = mg.batch_workload(12, seed=6, nmin=200, nmax=3000) + _edge_pairs()
    got, stats = _check_batch(pairs)
    assert stats["grid"] == 1
    assert sum(len(g["pairs"]) > 0 for g in got) >= 8
    assert stats["coplanar"] > 0  # the self pair


def test_batch_of_float_meshes_takes_the_wide_exact_stage():
    a = mg.jittered_icosphere(2, seed=1)
    b = mg.jittered_icosphere(2, seed=2, offset=(0.6, 0.1, 0.05))
    c = mg.jittered_icosphere(3, seed=3, offset=(-0.3, 0.5, 0.0))
    pairs = [(a, b), (b, c), (c, a), (a, a)]
    got, stats = _check_batch(pairs, rebuild_twice=False)
    assert stats["grid"] == 0 and stats["exact_pairs"] > 0


def test_batch_with_residual_pairs_merges_on_the_host():
    # tiny coordinates near x = 0 in one mesh: some pairs are left to the host
    # stage, merged with the GPU's and split per mesh pair there
    a = mg.jittered_icosphere(2, seed=1)
    b = mg.jittered_icosphere(2, seed=2, offset=(0.6, 0.1, 0.05))
    c = mg.jittered_icosphere(3, seed=3, offset=(-0.3, 0.5, 0.0))
    xa = a[0].copy()
    near = np.nonzero(np.abs(xa[:, 0]) < 0.05)[0]
    xa[near, 0] = np.ldexp(1.0, -80)
    pairs = [(b, c), ((xa, a[1]), b), (c, (xa, a[1])), (b, b)]
    got, stats = _check_batch(pairs, rebuild_twice=True)
    assert stats["host_pairs"] > 0 and stats["exact_pairs"] > 0


def test_batch_spanning_many_sort_tiles_and_chunks():
    # enough triangles for multi-tile sorts, multi-chunk builds and mesh
    # subtrees whose roots are top nodes
    pairs = mg.batch_workload(40, seed=8, nmin=2000, nmax=60000)
    _check_batch(pairs, rebuild_twice=True)


def test_batch_context_refuses_single_pair_query():
    pairs = mg.batch_workload(2, seed=6, nmin=200, nmax=400)
    eng = gi.Intersector()
    try:
        eng.set_batch(pairs)
        with pytest.raises(gi.GiError):
            eng.intersect()
        with pytest.raises(gi.GiError):
            eng.set_transform(1, [1.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0])
    finally:
        eng.close()
