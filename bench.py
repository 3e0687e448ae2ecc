"""Benchmark: exact tri-pairs/s for two 2^20-triangle meshes on one MI355X.

Workload (one step): a displaced UV sphere (1024x512, 2^20 triangles) and a
rotated torus (1024x512, 2^20 triangles), both snapped to the 2^24 integer
grid, resident in HBM before timing. Each step runs the whole device path from
those arrays: vertex preparation, FP32 boxes and Morton sort for both meshes,
the LBVH of mesh A (B's triangles in Morton order are the traversal's
queries, read with their boxes; nothing walks a tree of B), BVH traversal,
FP64 filter,
exact int128 stage with
exact segments, sort of the intersecting pairs and gather of their segments
(output left in device memory).

value = exact candidate pairs decided per second (candidates / step time),
inputs resident in HBM. `e2e` times the same work from pinned host arrays
(both meshes uploaded, A's BVH built, the query, sorted pairs and exact
segments copied to host memory), labelled separately.
roofline = HBM roofline of the dominant kernel, the longest of the step's
three largest single launches (staging of both meshes, A's BVH build and the
traversal), chosen in the warm-up steps, which time all three with the
library's HIP events; the timed steps bracket only that kernel, and the
others are reported under roofline_other from a few more steps that time all
three. Achieved bandwidth = the kernel's bytes per
launch / its average duration, where the bytes are the cost model's, capped
by the measured HBM traffic of the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
summary when that is lower (the traversal reads only the part of A's tree
near B); the uncapped figure is reported as achieved_model.
pipeline.predicates: orient3d / orient2d signs evaluated per cascade stage
(FP64 filter, GPU exact stage, host residual stage) and their rates.
Other legs: batched (4,096 pairs: one pass vs the per-pair loop, both from
device-resident inputs and both from host arrays), degeneracy (the 512x512
height field against its half-cell shift and its 90-degree copy), serving
(unsnapped, the reference's transform semantics, and snapped).
The CPU baseline is the oracle (C restatement of the reference CPU path:
box sweep + filtered exact predicates + exact segments) on the same meshes,
timed on this host's cores (median of 5 runs after a warm-up), plus one
thread on a sample.
"""
from __future__ import annotations

import argparse
import ctypes
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


def kernel_bytes(name, st, ntA, ntB, nvA, nvB):
    """Compulsory bytes of one launch of a single kernel of the step (one
    launch of each per step; A's tree is the only tree built: B's triangles in
    Morton order are the traversal's queries)."""
    if name == "k_stage_mesh":
        # both meshes in one launch, neither moved (identity transform, source
        # already on the integer grid): the launch writes no vertex copies and
        # reads each source vertex once through the triangles that use it, as
        # the 16-byte integer copy made when the mesh was first staged; per
        # triangle: indices 12 B read, box 32 B + Morton key and id 8 B written
        # (the sort's bucket counts are atomics on 16 KB)
        return 16.0 * (nvA + nvB) + 52.0 * (ntA + ntB)
    if name == "k_build_chunks":
        # per leaf of A: sorted key 4 B + triangle id 4 B + its FP32 box 32 B
        # read; leaf and internal node 32 B each + leaf range 8 B written
        n = float(ntA)
        return 4 * n + 4 * n + 32 * n + 32 * (2 * n - 1) + 8 * (n - 1)
    if name == "k_traverse":
        # the cost model's traversal read, A's nodes once (2 x 32 B per
        # triangle of A); B's triangles once (id 4 B + box 32 B); 8 B per
        # candidate pair written
        return 64.0 * ntA + 36.0 * ntB + 8 * st["traversal_hits"]
    raise KeyError(name)


def measured_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 PMC summary of this
    workload (FETCH_SIZE and WRITE_SIZE runs, each scaled by what the counter
    reads for a plain device copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_sphere_torus_1Mx1M_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


class PinnedArrays:
    """Host arrays in pinned (page-locked) memory from hipHostMalloc of the
    HIP runtime the library uses: the source of the end-to-end timing."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ptrs = []

    def copy_of(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipHostMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 1)), 0) == 0
        self.ptrs.append(p)
        out = np.frombuffer((ctypes.c_char * max(a.nbytes, 1)).from_address(p.value), dtype=a.dtype,
                            count=a.size).reshape(a.shape)
        out[...] = a
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipHostFree(p)
        self.ptrs = []


class DeviceArena:
    """Device copies of host arrays (hipMalloc of the library's HIP runtime),
    for queries from device-resident inputs."""

    def __init__(self, nbytes):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.base = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.base), ctypes.c_size_t(max(nbytes, 256))) == 0
        self.off = 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        ptr = self.base.value + self.off
        if a.nbytes:
            assert self.hip.hipMemcpy(ctypes.c_void_p(ptr), a.ctypes.data_as(ctypes.c_void_p),
                                      ctypes.c_size_t(a.nbytes), 1) == 0  # host to device
        self.off += (a.nbytes + 255) // 256 * 256
        return ptr

    def free(self):
        self.hip.hipFree(self.base)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def device_copy(nbytes, reps):
    """Plain device-to-device copies of nbytes (the reference dispatch for
    scaling the rocprofv3 FETCH_SIZE / WRITE_SIZE counters), through the HIP
    runtime the library itself uses."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    try:
        for _ in range(reps):
            assert hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0  # device to device
        assert hip.hipDeviceSynchronize() == 0
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)


def cpu_baseline(xa, ta, xb, tb, stride, reps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py

    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    sub = ta[::stride]
    times, r = [], None
    for i in range(reps + 1):  # one warm-up, then reps timed runs
        t0 = time.perf_counter()
        r = oracle_py.intersect(xa, sub, xb, tb, nthreads=cores, use_filter=True, want_segments=True,
                                decode_segments=False)
        if i:
            times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    c = r["stats"]["n_candidates"]
    frac = "all" if stride == 1 else f"every {stride}th triangle"
    # one thread on every 16th triangle of A (about as much work as one run above per core)
    sub1 = ta[::16 * stride]
    t1 = []
    for i in range(3):
        t0 = time.perf_counter()
        r1 = oracle_py.intersect(xa, sub1, xb, tb, nthreads=1, use_filter=True, want_segments=True,
                                 decode_segments=False)
        if i:
            t1.append(time.perf_counter() - t0)
    d1 = float(np.median(t1))
    c1 = r1["stats"]["n_candidates"]
    return {"value": c / dt, "unit": "exact tri-pairs/s", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{frac} of mesh A ({len(sub)} of {len(ta)}) vs all of mesh B, median of {reps} runs after "
                      f"one warm-up: {c} candidates, {r['stats']['n_pairs']} pairs with exact segments, "
                      f"{dt:.2f} s per run ({sum(times):.1f} s timed)",
            "single_thread": {"value": c1 / d1, "unit": "exact tri-pairs/s", "cores": 1,
                              "sample": f"every {16 * stride}th triangle of mesh A ({len(sub1)}) vs all of mesh B, "
                                        f"median of 2 runs after one warm-up: {c1} candidates, {d1:.2f} s per run"}}


def serving_leg(gi, mg, n_requests, qps, slots, snap):
    """Serving workload: a resident 2,097,152-triangle displaced sphere; Poisson
    requests (seeded) moving a 262,144-triangle torus (80 %) or a 4,096-triangle
    probe (20 %) by a random rigid motion; each request = transform + refit of
    the moving mesh, query against the resident BVH, exact segments, host copy.
    snap=False keeps the moved doubles (the reference's transform; the GPU
    exact stage works in per-pair integer frames), snap=True rounds them to the
    grid. Open loop at `qps`, then a closed burst of the same requests (all
    queued at once) for the saturation throughput."""
    import queue
    import threading

    w = mg.serving_workload(n_requests=n_requests, qps=qps)
    t_setup = time.perf_counter()
    srv = gi.Server(*w["static"], slots=slots, snap=snap, want_segments=True)
    for m in w["meshes"]:
        srv.add_mesh(*m)
    setup_s = time.perf_counter() - t_setup
    reqs = w["requests"]
    for mid in (0, 1):  # warm both meshes on every slot
        first = next(m for _, k, m in reqs if k == mid)
        for t in [srv.submit(mid, first) for _ in range(2 * slots)]:
            srv.wait(t, decode=False)

    def run(arrivals):
        todo, out = queue.Queue(), []

        def waiter():
            while True:
                t = todo.get()
                if t is None:
                    return
                r = srv.wait(t, decode=False)
                out.append((r["info"], len(r["pairs"]), r["n_segments"], r["stats"]))

        th = threading.Thread(target=waiter)
        th.start()
        lag = []
        t0 = time.perf_counter()
        for (_, mid, m), ta in zip(reqs, arrivals):
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            lag.append(time.perf_counter() - t0 - ta)
            todo.put(srv.submit(mid, m))
        todo.put(None)
        th.join()
        wall = time.perf_counter() - t0
        return out, wall, lag

    open_loop, wall_open, lag = run([a for a, _, _ in reqs])
    burst, wall_burst, _ = run([0.0] * len(reqs))
    srv.close()

    def pct(v, q):
        return float(np.percentile(np.asarray(v), q))

    lat = [i["latency_ms"] for i, _, _, _ in open_loop]
    svc = [i["service_ms"] for i, _, _, _ in open_loop]
    by_mesh = {}
    for (_, mid, _), (i, npairs, _, _) in zip(reqs, open_loop):
        by_mesh.setdefault(mid, []).append(i["service_ms"])
    return {
        "workload": f"static {w['static'][1].shape[0]} triangles resident; {len(reqs)} Poisson requests at "
                    f"{qps:g}/s (seed 7): torus {w['meshes'][0][1].shape[0]} / probe "
                    f"{w['meshes'][1][1].shape[0]} triangles, random rigid motion "
                    f"({'snapped to the grid' if snap else 'moved doubles kept'}), exact segments to host",
        "snap_to_grid": bool(snap),
        "host_pairs": int(sum(s["host_pairs"] for _, _, _, s in open_loop)),
        "exact_pairs": int(sum(s["exact_pairs"] for _, _, _, s in open_loop)),
        "slots": slots, "requests": len(reqs), "qps_offered": qps,
        "qps_achieved": len(reqs) / wall_open,
        "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "max_ms": max(lat),
        "service_p50_ms": pct(svc, 50), "service_p99_ms": pct(svc, 99),
        "service_p50_torus_ms": pct(by_mesh.get(0, [0.0]), 50),
        "service_p50_probe_ms": pct(by_mesh.get(1, [0.0]), 50),
        "saturation_rps": len(reqs) / wall_burst,
        "burst_p50_ms": pct([i["latency_ms"] for i, _, _, _ in burst], 50),
        "burst_p99_ms": pct([i["latency_ms"] for i, _, _, _ in burst], 99),
        "submit_lag_p99_ms": pct(lag, 99) * 1000.0,
        "pairs_mean": float(np.mean([n for _, n, _, _ in open_loop])),
        "setup_s": setup_s,
    }


def batch_leg(gi, mg, n_pairs, steps, autotune=True):
    """Batched workload: n_pairs independent mesh pairs (seed 6), each side an
    icosphere, displaced sphere or torus with N log-uniform in [500, 200000],
    B placed at U[0.3, 1.5] radii from A, 2^24 grid. Batched = one pass of
    launches over the whole batch (gi_intersect_batch); plain = the
    single-pair API once per pair (set both meshes + gi_intersect). Compared
    two ways, each like for like: both from device-resident source arrays
    (batched: a query with rebuild after gi_set_batch; plain:
    gi_set_mesh_device x2 + gi_intersect per pair, from one arena uploaded
    beforehand), and both end to end from host arrays (uploads included)."""
    from ctypes import byref

    t_gen = time.perf_counter()
    pairs = mg.batch_workload(n_pairs)
    gen_s = time.perf_counter() - t_gen
    ntri = int(sum(a[1].shape[0] + b[1].shape[0] for a, b in pairs))
    free = gi.lib().gi_result_free

    eng = gi.Intersector()
    t0 = time.perf_counter()
    eng.set_batch(pairs)
    set_s = time.perf_counter() - t0
    tuning = eng.autotune(reps=1, rebuild=True) if autotune else eng.get_tuning()
    for _ in range(2):  # warm-up: arenas grow to their final sizes here
        res, offs = eng.intersect_batch_raw(output_on_device=True, want_segments=True, rebuild=True)
        free(byref(res))
    t0 = time.perf_counter()
    for _ in range(steps):
        res, offs = eng.intersect_batch_raw(output_on_device=True, want_segments=True, rebuild=True)
        st = res.stats.as_dict()
        free(byref(res))
    step_s = (time.perf_counter() - t0) / steps
    counts = np.diff(offs.astype(np.int64))
    # end to end from host arrays: set + one query (pairs to device memory)
    t0 = time.perf_counter()
    eng.set_batch(pairs)
    res, _ = eng.intersect_batch_raw(output_on_device=True, want_segments=True, rebuild=False)
    free(byref(res))
    e2e_batch_s = time.perf_counter() - t0
    sample = {i: r for i, r in enumerate(eng.intersect_batch(want_segments=False)[0]) if i % 97 == 0}
    eng.close()

    # plain, from device-resident source arrays: every pair's arrays uploaded
    # to one device arena first (not timed)
    nbytes = sum(sum((a.nbytes + 255) // 256 * 256 for a in m) for pr in pairs for m in pr)
    arena = DeviceArena(nbytes)
    dev = [tuple((arena.put(x), x.shape[0], arena.put(t), t.shape[0]) for x, t in pr) for pr in pairs]
    plain = gi.Intersector()
    plain_counts, plain_cands, plain_sample = [], 0, {}
    for _ in range(2):  # warm-up: arenas grow
        for side, (px, nvx, pt, ntx) in enumerate(dev[0]):
            plain.set_mesh_device(side, px, nvx, pt, ntx)
        free(byref(plain.intersect_raw(output_on_device=True, want_segments=True)))
    t0 = time.perf_counter()
    for d in dev:
        for side, (px, nvx, pt, ntx) in enumerate(d):
            plain.set_mesh_device(side, px, nvx, pt, ntx)
        res = plain.intersect_raw(output_on_device=True, want_segments=True)
        plain_counts.append(int(res.n_pairs))
        plain_cands += int(res.stats.candidates)
        free(byref(res))
    plain_resident_s = time.perf_counter() - t0
    arena.free()
    # plain, end to end from host arrays
    plain_e2e_counts = []
    t0 = time.perf_counter()
    for (xa, ta), (xb, tb) in pairs:
        plain.set_mesh(0, xa, ta)
        plain.set_mesh(1, xb, tb)
        res = plain.intersect_raw(output_on_device=True, want_segments=True)
        plain_e2e_counts.append(int(res.n_pairs))
        free(byref(res))
    plain_s = time.perf_counter() - t0
    for i in sample:
        (xa, ta), (xb, tb) = pairs[i]
        plain.set_mesh(0, xa, ta)
        plain.set_mesh(1, xb, tb)
        plain_sample[i] = plain.intersect(want_segments=False)["pairs"]
    plain.close()
    parity = bool(np.array_equal(counts, np.asarray(plain_counts)) and plain_cands == st["candidates"] and
                  plain_e2e_counts == plain_counts and
                  all(np.array_equal(sample[i]["pairs"], plain_sample[i]) for i in sample))
    return {
        "workload": f"{n_pairs} mesh pairs (seed 6), {ntri} triangles in total, N log-uniform in [500, 200000], "
                    "icosphere / displaced sphere / torus, centre distance U[0.3, 1.5] radii, 2^24 grid",
        "pairs": n_pairs, "triangles": ntri,
        "batched_ms": step_s * 1e3,
        "batched_mesh_pairs_per_s": n_pairs / step_s,
        "batched_exact_tri_pairs_per_s": st["candidates"] / step_s,
        "batched_stage_ms": {k: st[k] for k in ("ms_build", "ms_traverse", "ms_filter", "ms_exact", "ms_output")},
        "plain_resident_ms": plain_resident_s * 1e3,
        "plain_resident_mesh_pairs_per_s": n_pairs / plain_resident_s,
        "speedup_resident": plain_resident_s / step_s,
        "e2e_batched_ms": e2e_batch_s * 1e3,
        "plain_e2e_ms": plain_s * 1e3,
        "speedup_e2e": plain_s / e2e_batch_s,
        "candidates": st["candidates"], "intersecting_pairs": st["pairs"], "segments": st["segments"],
        "parity_with_plain": parity,
        "tuning": tuning,
        "generate_s": gen_s, "set_batch_s": set_s,
    }


def degeneracy_leg(gi, mg, steps):
    """The 512 x 512 height field (524,288 triangles, integer coordinates,
    heights in {0..8}) against its (1/2, 1/2, 0)-cell shift and against its
    90-degree copy about z, where coplanar, shared-edge and zero-orientation
    contacts dominate. Each step as in the main leg (device-resident meshes,
    both restaged, A's BVH rebuilt, query with exact segments). Reports the
    fraction of candidates each cascade stage decided and the exact stage's
    time."""
    (p, t), (sft, _), (rot, _) = mg.height_field_pairs(cells=512)
    out = {}
    for name, q in (("shift", sft), ("rot90", rot)):
        eng = gi.Intersector()
        eng.set_mesh(0, p, t)
        eng.set_mesh(1, q, t)
        eng.build()

        def step():
            res = eng.intersect_raw(output_on_device=True, want_segments=True, rebuild=True)
            s = res.stats.as_dict()
            gi.lib().gi_result_free(gi.C.byref(res))
            return s

        for _ in range(2):
            step()
        t0 = time.perf_counter()
        sts = [step() for _ in range(steps)]
        ms = (time.perf_counter() - t0) * 1e3 / steps
        eng.close()
        s = sts[-1]
        c = max(s["candidates"], 1)
        out[name] = {
            "triangles": int(t.shape[0]), "ms_per_query": ms, "exact_tri_pairs_per_s": s["candidates"] / (ms / 1e3),
            "candidates": s["candidates"], "pairs": s["pairs"], "segments": s["segments"], "coplanar": s["coplanar"],
            "degenerate": s["degenerate"], "grid": s["grid"],
            "resolved": {"filter_no": s["filter_no"] / c, "filter_green": s["filter_green"] / c,
                         "exact_yellow": (s["filter_yellow"] - s["host_pairs"]) / c, "host": s["host_pairs"] / c},
            "exact_yellow_yes": s["exact_yellow_yes"], "exact_yellow_no": s["exact_yellow_no"],
            "predicates": {"filter": s["filter_predicates"], "exact": s["exact_predicates"],
                           "filter_per_s": s["filter_predicates"] / max(s["ms_filter"], 1e-6) * 1e3,
                           "exact_per_s": s["exact_predicates"] / max(s["ms_exact"], 1e-6) * 1e3},
            "stage_ms": {k: float(np.mean([x[k] for x in sts])) for k in
                         ("ms_build", "ms_traverse", "ms_filter", "ms_exact", "ms_output")},
        }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--res", type=int, default=1024, help="UV resolution (1024 -> 2^20 triangles per mesh)")
    ap.add_argument("--cpu-stride", type=int, default=1)
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--graph", action="store_true",
                    help="replay each step's launches up to the counter read-back as a HIP graph")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched measurement")
    ap.add_argument("--batch-pairs", type=int, default=4096)
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched measurement (for kernel profiles of it)")
    ap.add_argument("--no-degeneracy", action="store_true", help="skip the height-field degeneracy measurement")
    ap.add_argument("--degeneracy-steps", type=int, default=10)
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-qps", type=float, default=200.0)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--check", action="store_true", help="compare the pair set with the oracle (slow)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB (3 times): the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the library runs a query on "
              "one device (several GPUs would each take a slice of B's Morton-ordered leaves against a "
              "local copy of A's tree)", file=sys.stderr)
        sys.exit(2)

    gi = ge.package_module("gi_ctypes")
    mg = ge.package_module("meshgen")
    if args.batch_only:
        print(json.dumps({"batched": batch_leg(gi, mg, args.batch_pairs, args.batch_steps, not args.no_autotune)}),
              flush=True)
        return
    (xa, ta), (xb, tb) = mg.sphere_torus(args.res, args.res // 2)
    # both meshes are uploaded once: every step reads them from HBM
    eng = gi.Intersector(device=0)
    eng.set_mesh(0, xa, ta)
    eng.set_mesh(1, xb, tb)
    eng.build()

    tuning = eng.get_tuning() if args.no_autotune else eng.autotune(reps=3, rebuild=True)

    def step(timing, convert=True):
        res = eng.intersect_raw(output_on_device=True, want_segments=True, rebuild=True, use_graph=args.graph,
                                stage_timing=timing)
        st = type(res.stats).from_buffer_copy(res.stats)  # the C struct, copied
        gi.lib().gi_result_free(gi.C.byref(res))
        return st.as_dict() if convert else st

    # a step returns once its results are complete in device memory (the
    # library synchronises its streams), so the clock brackets finished work.
    # The warm-up steps time every stage; the largest single launches of a
    # step are the staging of both meshes (k_stage_mesh), A's BVH build
    # (k_build_chunks) and the traversal (k_traverse), and the longest of them
    # in the warm-up is the dominant kernel. The timed steps record HIP events
    # around that one launch only (each event costs the queue a few
    # microseconds); a few more steps with every stage timed give the stage
    # breakdown.
    flags = {"k_stage_mesh": gi.TIME_STAGING, "k_build_chunks": gi.TIME_BVH_CHUNKS, "k_traverse": gi.TIME_TRAVERSE}
    stat_of = {"k_stage_mesh": "ms_stage_meshes", "k_build_chunks": "ms_bvh_chunks", "k_traverse": "ms_traverse"}
    warm = [step(0) for _ in range(max(args.warmup, 1))]
    warm_ms = {k: float(np.mean([w[stat_of[k]] for w in warm])) for k in flags}
    dom = max(warm_ms, key=warm_ms.get)
    raw = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        raw.append(step(flags[dom], convert=False))  # the stats struct; its Python dict after the clock stops
    t1 = time.perf_counter()
    stats = [r.as_dict() for r in raw]
    ms = (t1 - t0) * 1000.0 / args.steps
    st = stats[-1]
    for k in ("candidates", "pairs", "segments"):
        assert all(s[k] == st[k] for s in stats), f"{k} varies between steps"
    staged = [step(0) for _ in range(5)]
    avg = {k: float(np.mean([s[k] for s in staged])) for k in
           ("ms_total", "ms_build", "ms_traverse", "ms_filter", "ms_exact", "ms_output", "ms_bvh_chunks")}
    dom_ms = float(np.mean([s[stat_of[dom]] for s in stats]))
    value = st["candidates"] / (ms / 1000.0)

    # the dominant kernel, timed by HIP events around its one launch per step
    # in the timed steps; the other two from the warm-up steps
    assert all(s["bvh_chunk_leaves"] in (0, ta.shape[0]) for s in stats + warm), "the build covers A's tree only"
    kernel_ms = dict(warm_ms)
    kernel_ms[dom] = dom_ms
    roofs = {}
    for k, kms in kernel_ms.items():
        model = kernel_bytes(k, st, ta.shape[0], tb.shape[0], xa.shape[0], xb.shape[0])
        traffic = measured_traffic(k)
        kb = min(model, traffic) if traffic else model
        achieved = kb / (kms / 1000.0) / 1e9
        roofs[k] = {"bound": "hbm", "kernel": k, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel_ms": kms,
                    "launches_timed": len(stats) if k == dom else len(warm),
                    "timed_in": "the timed steps" if k == dom else "the warm-up steps",
                    "algorithmic_bytes": kb, "model_bytes": model,
                    "achieved_model": model / (kms / 1000.0) / 1e9}
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
        "config": {"workload": f"sphere x torus, {ta.shape[0]} x {tb.shape[0]} triangles, 2^24 grid; per step "
                               "both meshes staged and Morton-sorted, A's BVH built (B's triangles in Morton "
                               "order are the queries), query with exact segments",
                   "triangles_a": int(ta.shape[0]), "triangles_b": int(tb.shape[0]),
                   "vertices_a": int(xa.shape[0]), "vertices_b": int(xb.shape[0])},
        "tuning": tuning,
        "graph_steps": sum(s["graph"] for s in stats),
        "roofline": roofs[dom],
        "roofline_other": [r for k, r in roofs.items() if k != dom],
        "timing_events": f"the timed steps record HIP events around {dom} only",
        "pipeline": {"candidates": st["candidates"], "traversal_hits": st["traversal_hits"], "pairs": st["pairs"],
                     "segments": st["segments"], "filter_no": st["filter_no"], "filter_green": st["filter_green"],
                     "filter_yellow": st["filter_yellow"], "exact_yellow_yes": st["exact_yellow_yes"],
                     "coplanar": st["coplanar"], "degenerate": st["degenerate"],
                     "intersecting_pairs_per_s": st["pairs"] / (ms / 1000.0),
                     "pair_space_per_s": ta.shape[0] * tb.shape[0] / (ms / 1000.0),
                     "stage_ms": avg, "stage_ms_from": "5 more steps with every stage timed (stage_timing=0)",
                     "end_to_end_algorithmic_GBs": b_alg / (ms / 1000.0) / 1e9,
                     "end_to_end_hbm_frac": b_alg / (ms / 1000.0) / 1e9 / HBM_PEAK_GBS,
                     "predicates": {
                         "filter": st["filter_predicates"], "exact": st["exact_predicates"],
                         "host": st["host_predicates"],
                         "filter_per_s": st["filter_predicates"] / (avg["ms_filter"] / 1000.0),
                         "exact_per_s": st["exact_predicates"] / max(avg["ms_exact"] / 1000.0, 1e-9),
                         "per_s": (st["filter_predicates"] + st["exact_predicates"]) / (ms / 1000.0)},
                     "resolved": {k: v / max(st["candidates"], 1) for k, v in (
                         ("filter_no", st["filter_no"]), ("filter_green", st["filter_green"]),
                         ("exact_yellow", st["filter_yellow"] - st["host_pairs"]), ("host", st["host_pairs"]))}},
    }
    # end to end from host arrays (both meshes uploaded, A's BVH built, the
    # query) to sorted pairs and exact segments in host memory: from pinned
    # arrays (the figure of the end-to-end definition) and from pageable ones
    pin = PinnedArrays()
    pinned = tuple(pin.copy_of(a) for a in (xa, ta, xb, tb))
    e2e = {}
    for label, (a0, a1, b0, b1) in (("pinned", pinned), ("pageable", (xa, ta, xb, tb))):
        times = []
        for _ in range(6):
            t0 = time.perf_counter()
            eng.set_mesh(0, a0, a1)
            eng.set_mesh(1, b0, b1)
            res = eng.intersect_raw(output_on_device=False, want_segments=True)
            times.append(time.perf_counter() - t0)
            assert res.n_pairs == st["pairs"]
            gi.lib().gi_result_free(gi.C.byref(res))
        e2e[label] = float(np.median(times[1:]))
    pin.free()
    line["e2e"] = {"value": st["candidates"] / e2e["pinned"], "unit": "exact tri-pairs/s",
                   "ms": e2e["pinned"] * 1e3, "pageable_ms": e2e["pageable"] * 1e3,
                   "what": "from pinned host arrays: both meshes uploaded, staged and sorted, A's BVH built, the "
                           "query, sorted pairs and exact segments copied to host memory (median of 5 after one)"}
    if args.check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        r = eng.intersect(want_segments=False)
        ref = oracle_py.intersect(xa, ta, xb, tb, nthreads=os.cpu_count(), want_segments=False)
        line["check"] = bool(np.array_equal(r["pairs"], ref["pairs"]))
    eng.close()
    if args.copy_calib:
        device_copy(args.copy_calib << 20, 3)
        print(f"copy calibration: {args.copy_calib} MiB read + {args.copy_calib} MiB written per dispatch",
              file=sys.stderr)
    if not args.no_batch:
        line["batched"] = batch_leg(gi, mg, args.batch_pairs, args.batch_steps, not args.no_autotune)
    if not args.no_degeneracy:
        line["degeneracy"] = degeneracy_leg(gi, mg, args.degeneracy_steps)
    if not args.no_serve:
        line["serving"] = serving_leg(gi, mg, args.serve_requests, args.serve_qps, args.serve_slots, snap=False)
        line["serving_snapped"] = serving_leg(gi, mg, args.serve_requests, args.serve_qps, args.serve_slots,
                                              snap=True)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(xa, ta, xb, tb, args.cpu_stride, args.cpu_reps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
