"""Shared helpers for the tests: imports of the product binding, the mesh
generators and the oracle."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import __graft_entry__ as ge  # noqa: E402
import oracle_py  # noqa: E402

gi = ge.package_module("gi_ctypes")
mg = ge.package_module("meshgen")


def run_gpu(xa, ta, xb, tb, want_segments=True, rebuild=False):
    eng = gi.Intersector(device=0)
    try:
        eng.set_mesh(0, xa, ta)
        eng.set_mesh(1, xb, tb)
        return eng.intersect(want_segments=want_segments, rebuild=rebuild)
    finally:
        eng.close()
