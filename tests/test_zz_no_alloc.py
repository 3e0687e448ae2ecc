"""gi_intersect_noalloc (the entry point that allocates nothing) against the
existing one-shot entry point gi_intersect_oneshot, to the same bar as every
parity test: identical sorted pairs, identical exact segments (as reduced
rationals), identical candidate counts. Several calls in a row on one GPU,
each later one smaller than the one before it, so that each call must work
from the scratch array the previous call left behind."""
import numpy as np
import pytest

from gi_test_util import gi, mg, oracle_py

pytestmark = pytest.mark.gpu

GI_ERR_TOO_LARGE = 7


def _same_as_oneshot(xa, ta, xb, tb, grid):
    rc, got, need = gi.intersect_noalloc(xa, ta, xb, tb)
    assert rc == 0, (rc, need)
    ref = gi.intersect_oneshot(xa, ta, xb, tb)
    np.testing.assert_array_equal(got["pairs"], ref["pairs"])
    assert got["segments"] == ref["segments"]
    for k in ("candidates", "traversal_hits", "pairs", "segments", "coplanar", "degenerate", "grid"):
        assert got["stats"][k] == ref["stats"][k], k
    assert got["stats"]["grid"] == grid
    return got


def test_calls_in_a_row_match_the_oneshot_entry_point():
    # larger, then smaller, then doubles (the wide exact stage), then a result
    # set past the single-pass output (multi-tile sort after the read-back)
    big = _same_as_oneshot(*mg.sphere_torus(160, 80)[0], *mg.sphere_torus(160, 80)[1], 1)
    small = _same_as_oneshot(*mg.sphere_torus(64, 32)[0], *mg.sphere_torus(64, 32)[1], 1)
    assert len(big["pairs"]) > len(small["pairs"]) > 10
    (xa, ta), (xb, tb) = mg.icosphere_pair(level=3)
    dbl = _same_as_oneshot(xa, ta, xb, tb, 0)
    assert dbl["exact_limbs"] == 4 and len(dbl["pairs"]) > 10
    (p, t), _, (r, _) = mg.height_field_pairs(cells=96)
    many = _same_as_oneshot(p, t, r, t, 1)
    assert len(many["pairs"]) > 16384 and many["stats"]["coplanar"] > 0
    # and the first, smaller one again against the oracle
    (xa, ta), (xb, tb) = mg.sphere_torus(64, 32)
    rc, again, _ = gi.intersect_noalloc(xa, ta, xb, tb)
    ref = oracle_py.intersect(xa, ta, xb, tb, nthreads=8)
    assert rc == 0
    np.testing.assert_array_equal(again["pairs"], ref["pairs"])
    assert again["segments"] == ref["segments"]


def test_small_caller_buffers_report_the_sizes_needed():
    (xa, ta), (xb, tb) = mg.sphere_torus(96, 48)
    rc, _, need = gi.intersect_noalloc(xa, ta, xb, tb, pairs_cap=4, seg_cap=4)
    assert rc == GI_ERR_TOO_LARGE and need[0] > 4
    rc, got, _ = gi.intersect_noalloc(xa, ta, xb, tb, pairs_cap=need[0], seg_cap=need[1])
    assert rc == 0 and len(got["pairs"]) == need[0]
    np.testing.assert_array_equal(got["pairs"], gi.intersect_oneshot(xa, ta, xb, tb)["pairs"])


def test_limits_are_refused():
    # more triangles than the stated maximum: refused before any GPU work
    n = (1 << 17) + 1
    x = np.zeros((3, 3))
    t = np.zeros((n, 3), dtype=np.uint32)
    rc, _, _ = gi.intersect_noalloc(x, t, x, t[:1])
    assert rc == GI_ERR_TOO_LARGE
    # pairs whose integer frames need the host residual stage are refused
    (xa, ta), (xb, tb) = mg.icosphere_pair(level=3)
    xa = xa.copy()
    near = np.nonzero(np.abs(xa[:, 1]) < 0.05)[0]
    xa[near, 1] = np.ldexp(1.0, -80)
    rc, _, _ = gi.intersect_noalloc(xa, ta, xb, tb)
    assert rc == GI_ERR_TOO_LARGE
    # an empty mesh gives no pairs
    rc, got, _ = gi.intersect_noalloc(xa, ta, xb, np.zeros((0, 3), dtype=np.uint32))
    assert rc == 0 and len(got["pairs"]) == 0
