"""Mesh input and result output (include/gi_meshio.h), host code only.

The OFF cases are written by hand with their expected triangles, rows and drop
lists worked out from the reference loader's rules (fan triangulation
(p0, pk, pk+1), a row listed once per fan triangle with a repeated vertex,
polygons of fewer than 3 vertices dropped, later duplicates of a vertex set
dropped; RobustOffLoader.h:142-178)."""
import csv
import os
import random

import numpy as np
import pytest

from gi_test_util import gi, mg

MIXED_OFF = """# hand-made soup
OFF
# counts follow
7 7 0
0 0 0
1 0 0
1 1 0
0 1 0
0 0 1   # trailing comment
1 0 1 255 0 0 255
2 2 2
4 0 1 2 3
3 0 0 1
3 2 0 1
5 0 4 4 5 6
2 0 1
3 6 5 0
3 1 4 5
"""


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_off_canonical_faces_rows_and_drops(tmp_path):
    m = gi.read_off(_write(tmp_path, "mixed.off", MIXED_OFF))
    np.testing.assert_array_equal(m["tri"], [[0, 1, 2], [0, 2, 3], [0, 4, 5], [0, 5, 6], [1, 4, 5]])
    np.testing.assert_array_equal(m["face_to_row"], [0, 0, 3, 3, 6])
    np.testing.assert_array_equal(m["dropped_degenerate"], [1, 3, 4])
    np.testing.assert_array_equal(m["dropped_duplicate"], [2, 5])
    assert m["extra_fan_triangles"] == 3
    assert (m["soup_points"], m["soup_polygons"]) == (7, 7)
    np.testing.assert_array_equal(m["xyz"][5], [1.0, 0.0, 1.0])  # colour values ignored
    np.testing.assert_array_equal(m["xyz"][6], [2.0, 2.0, 2.0])


def test_off_counts_on_header_line_and_colour_faces(tmp_path):
    text = "COFF 4 2 0\n0 0 0 1 1 1 1\n1 0 0 1 1 1 1\n0 1 0 1 1 1 1\n0 0 1 1 1 1 1\n3 0 1 2 10 20 30\n3 0 1 3\n"
    m = gi.read_off(_write(tmp_path, "c.off", text))
    np.testing.assert_array_equal(m["tri"], [[0, 1, 2], [0, 1, 3]])
    np.testing.assert_array_equal(m["face_to_row"], [0, 1])


@pytest.mark.parametrize("text", [
    "PLY\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n3 0 1 2\n",     # not OFF
    "OFF\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n3 0 1 3\n",      # index out of range
    "OFF\n3 1 0\n0 0 0\n1 0 0\n",                      # truncated vertex list
    "OFF\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n",               # missing faces
    "OFF\n3 1 0\n0 0 0\n1 0 x\n0 1 0\n3 0 1 2\n",      # bad number
    "OFF\n3 1 0\n0 0 0\n1 0 0\n0 1 0\n4 0 1 2\n",      # fewer indices than announced
    "OFF BINARY\n",
])
def test_off_rejects_malformed_files(tmp_path, text):
    with pytest.raises(gi.GiError) as e:
        gi.read_off(_write(tmp_path, "bad.off", text))
    assert e.value.status == 8


def test_off_missing_file():
    with pytest.raises(gi.GiError):
        gi.read_off("/nonexistent/mesh.off")


def test_off_round_trip_is_bit_exact(tmp_path):
    xyz, tri = mg.jittered_icosphere(3, seed=4)
    xyz = xyz * np.array([1.0, 1e-7, 3e5])  # values that need all 17 digits
    p = str(tmp_path / "ico.off")
    gi.write_off(p, xyz, tri)
    m = gi.read_off(p)
    assert m["xyz"].tobytes() == np.ascontiguousarray(xyz).tobytes()
    np.testing.assert_array_equal(m["tri"], tri)
    np.testing.assert_array_equal(m["face_to_row"], np.arange(len(tri)))
    assert len(m["dropped_degenerate"]) == 0 and len(m["dropped_duplicate"]) == 0


HEADER = ("FaceA,FaceB,VA0_Idx,VA1_Idx,VA2_Idx,VB0_Idx,VB1_Idx,VB2_Idx,"
          "A_V0_X,A_V0_Y,A_V0_Z,A_V1_X,A_V1_Y,A_V1_Z,A_V2_X,A_V2_Y,A_V2_Z,"
          "B_V0_X,B_V0_Y,B_V0_Z,B_V1_X,B_V1_Y,B_V1_Z,B_V2_X,B_V2_Y,B_V2_Z").split(",")


def _meshes():
    a = mg.jittered_icosphere(1, seed=1)
    b = mg.jittered_icosphere(1, seed=2, offset=(0.5, 0.0, 0.0))
    return a, b


def _check_record(row, fa, fb, a, b):
    (xa, ta), (xb, tb) = a, b
    assert int(row[0]) == fa and int(row[1]) == fb
    assert [int(v) for v in row[2:5]] == list(ta[fa]) and [int(v) for v in row[5:8]] == list(tb[fb])
    want = np.concatenate([xa[ta[fa]].ravel(), xb[tb[fb]].ravel()])
    assert [float(v) for v in row[8:26]] == list(want)  # 17 significant digits round-trip exactly


def test_pairs_csv_has_the_reference_columns(tmp_path):
    a, b = _meshes()
    pairs = np.array([[0, 3], [5, 1], [7, 7]], dtype=np.uint32)
    p = str(tmp_path / "run.csv")
    gi.write_pairs_csv(p, pairs, a, b)
    rows = list(csv.reader(open(p)))
    assert rows[0] == HEADER
    assert len(rows) == 4
    for r, (fa, fb) in zip(rows[1:], pairs):
        _check_record(r, fa, fb, a, b)
    with pytest.raises(gi.GiError):
        gi.write_pairs_csv(p, np.array([[999, 0]]), a, b)


def test_pairs_diff_csv_categories_and_filters(tmp_path):
    a, b = _meshes()
    p1 = np.array([[3, 1], [0, 2], [5, 5], [0, 2]], dtype=np.uint32)   # duplicate collapses
    p2 = np.array([[5, 5], [1, 1], [0, 2]], dtype=np.uint32)
    p = str(tmp_path / "diff.csv")
    gi.write_pairs_diff_csv(p, p1, 4, p2, 9, "ALL", a, b)
    rows = list(csv.reader(open(p)))
    assert rows[0] == ["DiffCategory"] + HEADER
    got = [(r[0], int(r[1]), int(r[2])) for r in rows[1:]]
    assert got == [("ONLY_IN_4", 3, 1), ("ONLY_IN_9", 1, 1), ("IN_BOTH", 0, 2), ("IN_BOTH", 5, 5)]
    for r in rows[1:]:
        _check_record(r[1:], int(r[1]), int(r[2]), a, b)
    gi.write_pairs_diff_csv(p, p1, 4, p2, 9, "both", a, b)
    assert [(r[0], r[1], r[2]) for r in list(csv.reader(open(p)))[1:]] == [("IN_BOTH", "0", "2"), ("IN_BOTH", "5", "5")]
    gi.write_pairs_diff_csv(p, p1, 4, p2, 9, "exclusive2", a, b)
    assert [r[0] for r in list(csv.reader(open(p)))[1:]] == ["ONLY_IN_9"]
    gi.write_pairs_diff_csv(p, p1, 4, p2, 9, "nothing", a, b)
    assert len(list(csv.reader(open(p)))) == 1


def _limbs(v, L):
    u = v % (1 << (64 * L))
    return [(u >> (64 * k)) & ((1 << 64) - 1) for k in range(L)]


def test_segment_exports_render_exact_integers(tmp_path):
    """A hand-built host result with L-limb two's complement numbers (both
    signs, zero, extremes) must come back as the same decimal integers."""
    rnd = random.Random(5)
    L = 3
    pairs = np.array([[2, 9], [4, 1]], dtype=np.uint32)
    offs = np.array([0, 2, 3], dtype=np.uint32)
    S = 3
    vals = []
    for s in range(S):
        row = [rnd.randrange(-(1 << 150), 1 << 150) for _ in range(8)]
        row[3] = rnd.randrange(1, 1 << 120)
        row[7] = rnd.randrange(1, 1 << 120)
        vals.append(row)
    vals[1][0] = 0
    vals[1][1] = -(1 << 191)
    vals[1][2] = (1 << 191) - 1
    exact = np.array([[_limbs(v, L) for v in row] for row in vals], dtype=np.uint64).ravel()
    f64 = np.array([[rnd.uniform(-1e9, 1e9) for _ in range(6)] for _ in range(S)])
    exp2 = np.array([0, -3, 7], dtype=np.int32)
    res = gi.GiResult()
    res.n_pairs = 2
    res.pairs = pairs.ctypes.data_as(gi.C.POINTER(gi.C.c_uint32))
    res.seg_offsets = offs.ctypes.data_as(gi.C.POINTER(gi.C.c_uint32))
    res.n_segments = S
    res.exact_limbs = L
    res.on_device = 0
    res.seg_exact = exact.ctypes.data_as(gi.C.POINTER(gi.C.c_uint64))
    res.seg_exp2 = exp2.ctypes.data_as(gi.C.POINTER(gi.C.c_int32))
    res.seg_f64 = f64.ctypes.data_as(gi.C.POINTER(gi.C.c_double))
    p = str(tmp_path / "seg.csv")
    gi.write_segments(res, p, "csv")
    rows = list(csv.reader(open(p)))
    assert rows[0][:4] == ["pair", "face_a", "face_b", "segment"] and rows[0][-1] == "exp2"
    want_pair = [(0, 2, 9, 0), (0, 2, 9, 1), (1, 4, 1, 0)]
    for s, r in enumerate(rows[1:]):
        assert tuple(int(v) for v in r[:4]) == want_pair[s]
        assert [float(v) for v in r[4:10]] == list(f64[s])
        assert [int(v) for v in r[10:18]] == vals[s]
        assert int(r[18]) == exp2[s]
    o = str(tmp_path / "seg.obj")
    gi.write_segments(res, o, "obj")
    lines = [ln.split() for ln in open(o) if not ln.startswith("#")]
    assert [ln[0] for ln in lines] == ["v"] * 6 + ["l"] * 3
    assert lines[6:] == [["l", "1", "2"], ["l", "3", "4"], ["l", "5", "6"]]
    res.on_device = 1
    with pytest.raises(gi.GiError):
        gi.write_segments(res, p, "csv")
    os.remove(p)
