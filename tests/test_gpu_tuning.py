"""Launch-shape tuning (gi_set_tuning / gi_autotune): every setting gives the
same exact result; the autotuner returns a valid setting."""
import numpy as np
import pytest

from gi_test_util import gi, mg

pytestmark = pytest.mark.gpu

SETTINGS = [dict(traverse_variant=v) for v in range(4)] + \
    [dict(build_chunk=256), dict(filter_blocks=7, exact_blocks=3), dict(filter_blocks=8192, exact_blocks=4096)]


def _same(a, b):
    np.testing.assert_array_equal(a["pairs"], b["pairs"])
    assert a["segments"] == b["segments"]
    for k in ("candidates", "traversal_hits", "coplanar", "degenerate"):
        assert a["stats"][k] == b["stats"][k], k


def test_every_setting_gives_the_same_result():
    (xa, ta), (xb, tb) = mg.sphere_torus(256, 128)
    eng = gi.Intersector()
    try:
        eng.set_mesh(0, xa, ta)
        eng.set_mesh(1, xb, tb)
        ref = eng.intersect(rebuild=True)
        default = eng.get_tuning()
        for kw in SETTINGS:
            eng.set_tuning(**default)
            eng.set_tuning(**kw)
            _same(eng.intersect(rebuild=True), ref)
    finally:
        eng.close()


def test_autotune_on_single_pair_and_batch():
    (xa, ta), (xb, tb) = mg.sphere_torus(128, 64)
    eng = gi.Intersector()
    try:
        eng.set_mesh(0, xa, ta)
        eng.set_mesh(1, xb, tb)
        ref = eng.intersect()
        t = eng.autotune(reps=2)
        assert 0 <= t["traverse_variant"] <= 3 and t["build_chunk"] in (256, 512)
        assert eng.get_tuning() == t
        _same(eng.intersect(rebuild=True), ref)
        pairs = mg.batch_workload(6, seed=4, nmin=300, nmax=4000)
        eng.set_batch(pairs)
        before, _ = eng.intersect_batch()
        eng.autotune(reps=1)
        after, _ = eng.intersect_batch(rebuild=True)
        for x, y in zip(before, after):
            np.testing.assert_array_equal(x["pairs"], y["pairs"])
            assert x["segments"] == y["segments"]
    finally:
        eng.close()


def test_invalid_tuning_is_rejected():
    eng = gi.Intersector()
    try:
        for kw in (dict(traverse_variant=4), dict(build_chunk=1024), dict(filter_blocks=0), dict(exact_blocks=70000)):
            with pytest.raises(gi.GiError):
                eng.set_tuning(**kw)
        assert eng.get_tuning()["build_chunk"] == 512
    finally:
        eng.close()
