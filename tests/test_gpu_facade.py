"""The C++ KernelBVHController facade (construct -> runIntersectionPipeline ->
setTransformBoth -> run) against the oracle on the same (moved) vertices."""
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

from gi_test_util import ROOT, mg, oracle_py
from test_gpu_parity import gi_apply_rigid

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "cpp", "facade_test")


def _write_mesh(path, xyz, tri):
    with open(path, "wb") as f:
        np.array([xyz.shape[0], tri.shape[0]], dtype=np.uint32).tofile(f)
        np.ascontiguousarray(xyz, dtype=np.float64).tofile(f)
        np.ascontiguousarray(tri, dtype=np.uint32).tofile(f)


def _parse(path):
    runs, cur = [], None
    center = None
    for line in open(path):
        t = line.split()
        if t[0] == "run":
            cur = {"pairs": [], "segs": {}, "candidates": int(t[3])}
            runs.append(cur)
        elif t[0] == "pair":
            cur["pairs"].append((int(t[1]), int(t[2])))
        elif t[0] == "seg":
            a, b = int(t[1]), int(t[2])
            n = [int(x) for x in t[3:11]]
            s = Fraction(2) ** int(t[11])
            p = tuple(Fraction(n[k], n[3]) * s for k in range(3))
            q = tuple(Fraction(n[4 + k], n[7]) * s for k in range(3))
            cur["segs"].setdefault((a, b), []).append((p, q) if p <= q else (q, p))
        elif t[0] == "centerB":
            center = [float(x) for x in t[1:4]]
    return runs, center


@pytest.mark.parametrize("snap", [1, 0])
def test_facade_construct_run_transform_run(tmp_path, snap):
    """snap=1 rounds the moved vertices (grid exact stage); snap=0 is the
    reference's own behaviour, the moved doubles as they are (the per-pair
    integer frames of the wide exact stage), and also runs the per-call
    free function with the same motion."""
    assert os.path.exists(BIN), "tests/cpp/facade_test is built by __graft_entry__.build()"
    (xa, ta), (xb, tb) = mg.sphere_torus(96, 48)
    _write_mesh(tmp_path / "a.bin", xa, ta)
    _write_mesh(tmp_path / "b.bin", xb, tb)
    rot, tr = [12.0, -7.0, 30.0], [1.5e5, 2.0e4, -3.0e4]
    out = tmp_path / "out.txt"
    subprocess.run([BIN, str(tmp_path / "a.bin"), str(tmp_path / "b.bin"), str(out)] +
                   [str(x) for x in rot + tr] + [str(snap)], check=True, timeout=120)
    runs, center = _parse(out)
    assert len(runs) == (2 if snap else 3)
    np.testing.assert_allclose(center, xb.mean(axis=0), rtol=1e-12)
    xb_moved = gi_apply_rigid(xb, rot, tr, center)
    if snap:
        xb_moved = np.rint(xb_moved)
    else:
        assert not np.array_equal(xb_moved, np.rint(xb_moved))  # the moved mesh is off the grid
    for run, xbb in zip(runs, (xb, xb_moved, xb_moved)):
        ref = oracle_py.intersect(xa, ta, xbb, tb, nthreads=8)
        assert run["pairs"] == [tuple(p) for p in ref["pairs"].tolist()]
        assert run["candidates"] == ref["stats"]["n_candidates"]
        if not run["segs"]:  # the free function returns pairs only
            continue
        for i, (a, b) in enumerate(run["pairs"]):
            assert sorted(run["segs"][(a, b)]) == ref["segments"][i]
    assert runs[0]["pairs"] != runs[1]["pairs"]
