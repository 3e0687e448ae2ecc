"""Full-size parity of the batched and serving workloads (the bench's own
configurations), against the oracle.

Batching: the 4,096-pair batch (seed 6, triangle counts log-uniform in
[500, 200000] per mesh, 2.75e8 triangles in all) in one gi_intersect_batch;
a stratified sample of its pairs, several from every size decade of the
larger mesh, is checked pair for pair and segment for segment against the
oracle on that pair's arrays (a pair's result depends only on its own two
meshes), and every pair's count against the total.

Serving: the resident 2,097,152-triangle displaced sphere with the
262,144-triangle torus under two of the workload's seeded rigid motions and
the 4,096-triangle probe under one, through gi_server: the moved doubles kept
(the reference's transform), and one torus request snapped to the grid. Each
result equals the oracle on the host-moved vertices.

Both stop at the first pair that differs."""
import numpy as np
import pytest

from gi_test_util import gi, mg, oracle_py

pytestmark = pytest.mark.gpu


def test_full_batch_stratified_sample_matches_oracle():
    pairs = mg.batch_workload(4096)
    eng = gi.Intersector()
    try:
        eng.set_batch(pairs)
        res, offs = eng.intersect_batch_raw(output_on_device=False, want_segments=True)
        try:
            whole = gi.result_to_py(res, True, decode=False)
        finally:
            gi.lib().gi_result_free(gi.C.byref(res))
    finally:
        eng.close()
    offs = offs.astype(np.int64)
    assert offs[-1] == len(whole["pairs"]) == whole["stats"]["pairs"]
    assert whole["stats"]["grid"] == 1
    sizes = np.array([max(a[1].shape[0], b[1].shape[0]) for a, b in pairs])
    rng = np.random.default_rng(6)
    picked = []
    for lo, hi in ((0, 5000), (5000, 50000), (50000, 10 ** 9)):
        idx = np.nonzero((sizes >= lo) & (sizes < hi))[0]
        assert len(idx) > 0
        picked += list(rng.choice(idx, size=min(4, len(idx)), replace=False))
    so = whole["seg_offsets"].astype(np.int64)
    checked = 0
    for i in sorted(picked):
        (xa, ta), (xb, tb) = pairs[i]
        lo, hi = offs[i], offs[i + 1]
        ref = oracle_py.intersect(xa, ta, xb, tb, nthreads=16)
        np.testing.assert_array_equal(whole["pairs"][lo:hi], ref["pairs"], err_msg=f"pair {i}")
        segs = gi.decode_segments(whole["seg_exact"][so[lo]:so[hi]], None, so[lo:hi + 1] - so[lo])
        assert segs == ref["segments"], f"pair {i}"
        checked += len(ref["pairs"])
    assert checked > 1000  # the sample reaches intersecting pairs in number


def test_full_serving_requests_match_oracle():
    w = mg.serving_workload(n_requests=40)
    reqs = w["requests"]
    torus = [m for _, k, m in reqs if k == 0][:2]
    probe = [m for _, k, m in reqs if k == 1][:1]
    assert len(torus) == 2 and len(probe) == 1
    xs, ts = w["static"]
    assert ts.shape[0] == 2097152 and w["meshes"][0][1].shape[0] == 262144
    for snap, jobs in ((False, [(0, m) for m in torus] + [(1, m) for m in probe]), (True, [(0, torus[0])])):
        srv = gi.Server(xs, ts, slots=2, snap=snap, want_segments=True)
        try:
            for m in w["meshes"]:
                srv.add_mesh(*m)
            tickets = [srv.submit(mid, m) for mid, m in jobs]
            got = [srv.wait(t) for t in tickets]
        finally:
            srv.close()
        for (mid, m), r in zip(jobs, got):
            xm, tm = w["meshes"][mid]
            ref = oracle_py.intersect(xs, ts, mg.apply_affine(m, xm, snap_to_grid=snap), tm, nthreads=16)
            np.testing.assert_array_equal(r["pairs"], ref["pairs"], err_msg=f"mesh {mid} snap {snap}")
            assert r["segments"] == ref["segments"], f"mesh {mid} snap {snap}"
            assert r["stats"]["grid"] == int(snap)
            assert r["stats"]["host_pairs"] == 0
            assert len(r["pairs"]) > 0
