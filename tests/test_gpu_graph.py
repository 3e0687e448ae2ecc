"""HIP graph replay of a query (gi_options.use_graph): the same sequence of
rigid transforms on a graph context and an eager context must give identical
results at every step, including steps whose arenas grow (re-capture)."""
import numpy as np
import pytest

from gi_test_util import gi, mg

pytestmark = pytest.mark.gpu


def _transforms(n, seed=3):
    c = np.full(3, 2.0 * mg.GRID_SCALE)
    q = np.stack([mg.gaussian(seed, n, k) for k in range(4)], axis=1)
    t = np.stack([mg.uniform01(seed, n, 10 + k) for k in range(3)], axis=1) - 0.5
    return [mg.affine_about(mg.quat_matrix(q[i]), c, 0.3 * t[i] * mg.GRID_SCALE) for i in range(n)]


@pytest.mark.parametrize("reserve,snap", [(0, True), (1024, True), (0, False)])
def test_graph_replay_equals_eager(reserve, snap):
    (xa, ta), (xb, tb) = mg.sphere_torus(128, 64)
    eager = gi.Intersector(candidate_reserve=reserve)
    graph = gi.Intersector(candidate_reserve=reserve)
    try:
        for e in (eager, graph):
            e.set_mesh(0, xa, ta)
            e.set_mesh(1, xb, tb)
        replays = 0
        for i, m in enumerate(_transforms(9)):
            for e in (eager, graph):
                e.set_transform_matrix(1, m, snap=snap)
            r0 = eager.intersect()
            r1 = graph.intersect(use_graph=True)
            np.testing.assert_array_equal(r0["pairs"], r1["pairs"], err_msg=f"step {i}")
            assert r0["segments"] == r1["segments"], f"step {i}"
            for k in ("candidates", "traversal_hits", "pairs", "segments", "coplanar", "degenerate", "grid"):
                assert r0["stats"][k] == r1["stats"][k], (i, k)
            if r1["stats"]["graph"]:
                replays += 1
                assert r1["stats"]["ms_build"] > 0.0  # the graph's own event nodes time the stages
            assert r0["stats"]["graph"] == 0 and r0["stats"]["ms_build"] > 0.0
        assert replays >= 3
        assert sum(len(r["pairs"]) for r in (r0, r1)) > 0
    finally:
        eager.close()
        graph.close()


def test_graph_replay_with_rebuild_every_step():
    (xa, ta), (xb, tb) = mg.sphere_torus(96, 48)
    eng = gi.Intersector()
    try:
        eng.set_mesh(0, xa, ta)
        eng.set_mesh(1, xb, tb)
        ref = eng.intersect(rebuild=True)
        for _ in range(4):
            r = eng.intersect(rebuild=True, use_graph=True)
            np.testing.assert_array_equal(r["pairs"], ref["pairs"])
            assert r["segments"] == ref["segments"]
    finally:
        eng.close()


def test_graph_replays_with_split_bucket_sorts():
    # 2^20 triangles a side: the bucket sort splits the sphere's pole cells
    # over slice workgroups, whose slots each sort must leave cleared; graph
    # replays that rebuild, with eager queries between them, must match the
    # eager result and leave the same sorted keys
    (xa, ta), (xb, tb) = mg.sphere_torus(1024, 512)
    eng = gi.Intersector()
    try:
        eng.set_mesh(0, xa, ta)
        eng.set_mesh(1, xb, tb)
        ref = eng.intersect(rebuild=True)
        keys = eng.get_tree(0, len(ta), tree=False)["keys"]
        replays = 0
        for i in range(4):
            r = eng.intersect(rebuild=True, use_graph=True)
            replays += r["stats"]["graph"]
            np.testing.assert_array_equal(r["pairs"], ref["pairs"], err_msg=f"replay {i}")
            assert r["segments"] == ref["segments"], f"replay {i}"
            np.testing.assert_array_equal(eng.get_tree(0, len(ta), tree=False)["keys"], keys)
            e = eng.intersect(rebuild=True)
            np.testing.assert_array_equal(e["pairs"], ref["pairs"], err_msg=f"eager after replay {i}")
        assert replays >= 2 and len(ref["pairs"]) > 10000
    finally:
        eng.close()


def test_transform_switched_back_between_eager_queries_and_replays():
    """The device copy of a slot's transform is only refreshed when the staged
    transform changes. Switching between two transforms, eager and replayed
    queries interleaved (a replay writes the transform itself), every query
    must equal a fresh context's result for its transform; step 9 is an eager
    query with the transform of step 7 right after a replay with the other."""
    (xa, ta), (xb, tb) = mg.sphere_torus(96, 48)
    ms = _transforms(2, seed=5)
    refs = []
    for m in ms:
        e = gi.Intersector()
        try:
            e.set_mesh(0, xa, ta)
            e.set_mesh(1, xb, tb)
            e.set_transform_matrix(1, m, snap=True)
            refs.append(e.intersect())
        finally:
            e.close()
    assert not np.array_equal(refs[0]["pairs"], refs[1]["pairs"])
    seq = [(0, 0), (1, 0), (0, 0), (0, 1), (1, 1), (0, 1), (1, 1), (0, 0), (1, 1), (0, 0), (0, 0), (1, 0)]
    eng = gi.Intersector()
    try:
        eng.set_mesh(0, xa, ta)
        eng.set_mesh(1, xb, tb)
        for i, (k, g) in enumerate(seq):
            eng.set_transform_matrix(1, ms[k], snap=True)
            r = eng.intersect(use_graph=bool(g))
            np.testing.assert_array_equal(r["pairs"], refs[k]["pairs"], err_msg=f"step {i}")
            assert r["segments"] == refs[k]["segments"], f"step {i}"
            if i == 8:
                assert r["stats"]["graph"] == 1  # the step before the check of step 9 was a replay
    finally:
        eng.close()
