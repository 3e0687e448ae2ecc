"""OFF in -> GPU query -> the reference's CSV and the exact segment export
out, end to end through the C ABI."""
import csv
from fractions import Fraction

import numpy as np
import pytest

from gi_test_util import gi, mg, oracle_py

pytestmark = pytest.mark.gpu


def test_off_to_exports(tmp_path):
    (xa, ta), (xb, tb) = mg.sphere_torus(64, 32)
    pa, pb = str(tmp_path / "a.off"), str(tmp_path / "b.off")
    gi.write_off(pa, xa, ta)
    gi.write_off(pb, xb, tb)
    A, B = gi.read_off(pa), gi.read_off(pb)
    eng = gi.Intersector()
    try:
        eng.set_mesh(0, A["xyz"], A["tri"])
        eng.set_mesh(1, B["xyz"], B["tri"])
        res = eng.intersect_raw(output_on_device=False, want_segments=True)
        try:
            py = gi.result_to_py(res, True, True)
            seg_csv = str(tmp_path / "seg.csv")
            gi.write_segments(res, seg_csv, "csv")
            gi.write_segments(res, str(tmp_path / "seg.obj"), "obj")
        finally:
            gi.lib().gi_result_free(gi.C.byref(res))
    finally:
        eng.close()
    # the UV sphere's pole bands repeat the pole vertex in one triangle per
    # longitude: the loader drops those rows, so faces are renumbered
    deg = np.flatnonzero((ta[:, 0] == ta[:, 1]) | (ta[:, 1] == ta[:, 2]) | (ta[:, 0] == ta[:, 2]))
    np.testing.assert_array_equal(A["dropped_degenerate"], deg)
    np.testing.assert_array_equal(A["face_to_row"], np.setdiff1d(np.arange(len(ta)), deg))
    np.testing.assert_array_equal(A["tri"], ta[A["face_to_row"]])
    ref = oracle_py.intersect(A["xyz"], A["tri"], B["xyz"], B["tri"], nthreads=8)
    np.testing.assert_array_equal(py["pairs"], ref["pairs"])
    rows = list(csv.reader(open(seg_csv)))[1:]
    assert len(rows) == py["n_segments"] > 0
    per_pair = {}
    for r in rows:
        i = int(r[0])
        assert (int(r[1]), int(r[2])) == tuple(py["pairs"][i])
        X = [int(v) for v in r[10:18]]
        scale = Fraction(2) ** int(r[18])
        p, q = (tuple(Fraction(X[4 * e + c], X[4 * e + 3]) * scale for c in range(3)) for e in range(2))
        per_pair.setdefault(i, []).append((p, q) if p <= q else (q, p))
    for i, segs in enumerate(py["segments"]):
        assert sorted(per_pair.get(i, [])) == segs
    run_csv = str(tmp_path / "run.csv")
    gi.write_pairs_csv(run_csv, py["pairs"], (A["xyz"], A["tri"]), (B["xyz"], B["tri"]))
    assert len(list(csv.reader(open(run_csv)))) == len(py["pairs"]) + 1
