"""Serving loop (gi_server_*): one resident static mesh, moving meshes under
per-request rigid transforms. Each result must be identical whether the
request runs alone or in a mix with others on concurrent slots, and equal to
the oracle on the host-moved vertices."""
import numpy as np
import pytest

from gi_test_util import gi, mg, oracle_py

pytestmark = pytest.mark.gpu

UINT64_MAX = (1 << 64) - 1


def _small_workload(n=16):
    return mg.serving_workload(n_requests=n, static_res=(96, 48), torus_res=(48, 24), probe_res=(16, 8))


def _same(r1, r2):
    np.testing.assert_array_equal(r1["pairs"], r2["pairs"])
    assert r1["segments"] == r2["segments"]


def _server(w, **kw):
    srv = gi.Server(*w["static"], **kw)
    ids = [srv.add_mesh(*m) for m in w["meshes"]]
    assert ids == [0, 1]
    return srv


def test_results_alone_and_in_the_mix_are_identical_and_match_oracle():
    w = _small_workload()
    srv = _server(w, slots=4)
    try:
        alone = [srv.wait(srv.submit(mid, m)) for _, mid, m in w["requests"]]
        tickets = [srv.submit(mid, m) for _, mid, m in w["requests"]]
        mixed = [srv.wait(t) for t in tickets]
    finally:
        srv.close()
    for a, b in zip(alone, mixed):
        _same(a, b)
    assert len({r["info"]["slot"] for r in mixed}) > 1
    assert sum(len(r["pairs"]) > 0 for r in alone) >= len(alone) // 2
    xs, ts = w["static"]
    seen = set()
    for (_, mid, m), r in zip(w["requests"], alone):
        if mid in seen:
            continue
        seen.add(mid)
        xm, tm = w["meshes"][mid]
        ref = oracle_py.intersect(xs, ts, mg.apply_affine(m, xm), tm, nthreads=8)
        np.testing.assert_array_equal(r["pairs"], ref["pairs"])
        assert r["segments"] == ref["segments"]
        assert r["stats"]["grid"] == 1
    assert seen == {0, 1}


def test_unsnapped_requests_run_the_wide_exact_stage_and_match_oracle():
    # the moved vertices keep their doubles (the reference's transform): every
    # request's pairs and segments equal the oracle on the host-moved doubles,
    # alone and in the mix
    w = _small_workload(8)
    srv = _server(w, slots=4, snap=False)
    try:
        alone = [srv.wait(srv.submit(mid, m)) for _, mid, m in w["requests"]]
        tickets = [srv.submit(mid, m) for _, mid, m in w["requests"]]
        mixed = [srv.wait(t) for t in tickets]
    finally:
        srv.close()
    xs, ts = w["static"]
    for (_, mid, m), a, b in zip(w["requests"], alone, mixed):
        _same(a, b)
        xm, tm = w["meshes"][mid]
        ref = oracle_py.intersect(xs, ts, mg.apply_affine(m, xm, snap_to_grid=False), tm, nthreads=8)
        np.testing.assert_array_equal(a["pairs"], ref["pairs"])
        assert a["segments"] == ref["segments"]
        assert a["stats"]["grid"] == 0
    assert sum(len(r["pairs"]) > 0 for r in alone) >= len(alone) // 2


def test_admission_below_reserve_serialises_requests():
    w = _small_workload(8)
    srv = _server(w, slots=4)
    try:
        ref = [srv.wait(srv.submit(mid, m)) for _, mid, m in w["requests"]]
    finally:
        srv.close()
    # free HBM is never above a reserve of 2^64-1: a request only runs when
    # nothing else does
    srv = _server(w, slots=4, hbm_reserve_bytes=UINT64_MAX)
    try:
        tickets = [srv.submit(mid, m) for _, mid, m in w["requests"]]
        got = [srv.wait(t) for t in tickets]
    finally:
        srv.close()
    for a, b in zip(ref, got):
        _same(a, b)
    spans = sorted((r["info"]["t_admit_ms"], r["info"]["t_done_ms"]) for r in got)
    for (a0, d0), (a1, d1) in zip(spans, spans[1:]):
        assert a1 >= d0
    assert sum(r["info"]["refusals"] for r in got) > 0


def test_unknown_mesh_and_ticket_are_rejected():
    w = _small_workload(1)
    srv = _server(w, slots=1)
    try:
        with pytest.raises(gi.GiError):
            srv.submit(5, np.eye(3, 4))
        with pytest.raises(gi.GiError):
            srv.wait(12345)
    finally:
        srv.close()


def test_mesh_b_shared_into_an_a_slot_gets_its_whole_tree():
    # a B slot builds no BVH; shared into another context's A slot, whose tree
    # the traversal walks, it must build its tree first
    (xa, ta), (xb, tb) = mg.sphere_torus(128, 64)
    owner = gi.Intersector(device=0)
    user = gi.Intersector(device=0)
    try:
        owner.set_mesh(0, xb, tb)
        owner.set_mesh(1, xa, ta)
        first = owner.intersect()  # B = the sphere, no tree
        assert gi.lib().gi_share_mesh(user._ctx, 0, owner._ctx, 1) == 0
        user.set_mesh(1, xb, tb)
        got = user.intersect()
    finally:
        user.close()
        owner.close()
    ref = oracle_py.intersect(xa, ta, xb, tb, nthreads=8)
    np.testing.assert_array_equal(got["pairs"], ref["pairs"])
    assert got["segments"] == ref["segments"]
    assert len(first["pairs"]) == len(ref["pairs"])
