"""Benchmark: exact tri-pairs/s for two 2^20-triangle meshes on one MI355X.

Workload (one step): a displaced UV sphere (1024x512, 2^20 triangles) and a
rotated torus (1024x512, 2^20 triangles), both snapped to the 2^24 integer
grid, resident in HBM before timing. Each step runs the whole device path from
those arrays: vertex preparation and FP32 boxes, Morton sort + LBVH build +
refit for BOTH meshes, BVH traversal, FP64 filter, exact int128 stage with
exact segments, sort of the intersecting pairs and gather of their segments
(output left in device memory).

value = exact candidate pairs decided per second (candidates / step time).
The CPU baseline is the oracle (C restatement of the reference CPU path:
box sweep + filtered exact predicates + exact segments) on a strided sample
of mesh A against all of mesh B, timed on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import __graft_entry__ as ge  # noqa: E402

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(nvA, ntA, nvB, ntB, st):
    """Compulsory HBM bytes of one step (all stages), model:
    per mesh 12 B/vertex + 204 B/triangle (indices, boxes, Morton key/value,
    two 32-B BVH node writes and reads), per candidate 112 B, per pair 8 B,
    per exact segment 128 B."""
    per_mesh = 12 * (nvA + nvB) + 204 * (ntA + ntB)
    return per_mesh + 112 * st["candidates"] + 8 * st["pairs"] + 128 * st["segments"]


def kernel_bytes(name, ntA, ntB, st):
    """Compulsory bytes of one launch of a single kernel."""
    if name == "k_traverse":
        # every A node once (32 B), every B leaf once (32 B), 8 B per emitted pair
        return 32 * (2 * ntA - 1) + 32 * ntB + 8 * st["traversal_hits"]
    if name == "k_filter":
        # per emitted pair: 8 B read, two triangles (12 B indices + 3 x 16 B
        # int4 vertices); 8 B per pair forwarded to the exact stage
        return 128 * st["traversal_hits"] + 8 * (st["filter_green"] + st["filter_yellow"])
    if name == "k_exact":
        return 128 * st["exact_pairs"] + 16 * st["pairs"] + 128 * st["segments"]
    raise KeyError(name)


def cpu_baseline(xa, ta, xb, tb, stride):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py

    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    sub = ta[::stride]
    t0 = time.perf_counter()
    r = oracle_py.intersect(xa, sub, xb, tb, nthreads=cores, use_filter=True, want_segments=True,
                            decode_segments=False)
    dt = time.perf_counter() - t0
    c = r["stats"]["n_candidates"]
    return {"value": c / dt, "unit": "exact tri-pairs/s", "cores": cores, "kind": "port",
            "sample": f"every {stride}th triangle of mesh A ({len(sub)} of {len(ta)}) vs all of mesh B: "
                      f"{c} candidates, {r['stats']['n_pairs']} pairs, {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--res", type=int, default=1024, help="UV resolution (1024 -> 2^20 triangles per mesh)")
    ap.add_argument("--cpu-stride", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--check", action="store_true", help="compare the pair set with the oracle (slow)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB (3 times): the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this build "
              "(the multi-GPU design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch

    gi = ge.package_module("gi_ctypes")
    mg = ge.package_module("meshgen")
    (xa, ta), (xb, tb) = mg.sphere_torus(args.res, args.res // 2)
    dev = torch.device("cuda:0")
    dxa = torch.from_numpy(np.ascontiguousarray(xa)).to(dev)
    dta = torch.from_numpy(np.ascontiguousarray(ta).view(np.int32)).to(dev)
    dxb = torch.from_numpy(np.ascontiguousarray(xb)).to(dev)
    dtb = torch.from_numpy(np.ascontiguousarray(tb).view(np.int32)).to(dev)
    torch.cuda.synchronize()

    eng = gi.Intersector(device=0)
    eng.set_mesh_device(0, dxa.data_ptr(), xa.shape[0], dta.data_ptr(), ta.shape[0])
    eng.set_mesh_device(1, dxb.data_ptr(), xb.shape[0], dtb.data_ptr(), tb.shape[0])

    def step():
        res = eng.intersect_raw(output_on_device=True, want_segments=True, rebuild=True)
        st = res.stats.as_dict()
        gi.lib().gi_result_free(gi.C.byref(res))
        return st

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(step())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1000.0 / args.steps
    st = stats[-1]
    for k in ("candidates", "pairs", "segments"):
        assert all(s[k] == st[k] for s in stats), f"{k} varies between steps"
    avg = {k: float(np.mean([s[k] for s in stats])) for k in
           ("ms_total", "ms_build", "ms_traverse", "ms_filter", "ms_exact", "ms_output")}
    value = st["candidates"] / (ms / 1000.0)

    # dominant single kernel among the timed stages that are one kernel each
    single = {"k_traverse": avg["ms_traverse"], "k_filter": avg["ms_filter"], "k_exact": avg["ms_exact"]}
    dom = max(single, key=single.get)
    kb = kernel_bytes(dom, ta.shape[0], tb.shape[0], st)
    achieved = kb / (single[dom] / 1000.0) / 1e9
    b_alg = algorithmic_bytes(xa.shape[0], ta.shape[0], xb.shape[0], tb.shape[0], st)

    line = {
        "metric": "exact tri-pairs/s at 1Mx1M",
        "value": value,
        "unit": "exact tri-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64+int128",
        "data": "synthetic (seeded displaced sphere and rotated torus, snapped to the 2^24 grid)",
        "config": {"workload": f"sphere x torus, {ta.shape[0]} x {tb.shape[0]} triangles, 2^24 grid, "
                               "both BVH builds + query + exact segments per step",
                   "triangles_a": int(ta.shape[0]), "triangles_b": int(tb.shape[0]),
                   "vertices_a": int(xa.shape[0]), "vertices_b": int(xb.shape[0])},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel_ms": single[dom], "algorithmic_bytes": kb},
        "pipeline": {"candidates": st["candidates"], "traversal_hits": st["traversal_hits"], "pairs": st["pairs"],
                     "segments": st["segments"], "filter_no": st["filter_no"], "filter_green": st["filter_green"],
                     "filter_yellow": st["filter_yellow"], "exact_yellow_yes": st["exact_yellow_yes"],
                     "coplanar": st["coplanar"], "degenerate": st["degenerate"],
                     "intersecting_pairs_per_s": st["pairs"] / (ms / 1000.0),
                     "pair_space_per_s": ta.shape[0] * tb.shape[0] / (ms / 1000.0),
                     "stage_ms": avg, "end_to_end_algorithmic_GBs": b_alg / (ms / 1000.0) / 1e9,
                     "end_to_end_hbm_frac": b_alg / (ms / 1000.0) / 1e9 / HBM_PEAK_GBS},
    }
    if args.check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        r = eng.intersect(want_segments=False)
        ref = oracle_py.intersect(xa, ta, xb, tb, nthreads=os.cpu_count(), want_segments=False)
        line["check"] = bool(np.array_equal(r["pairs"], ref["pairs"]))
    eng.close()
    if args.copy_calib:
        src = torch.empty(args.copy_calib * (1 << 20) // 4, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copy calibration: {args.copy_calib} MiB read + {args.copy_calib} MiB written per dispatch",
              file=sys.stderr)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(xa, ta, xb, tb, args.cpu_stride)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
