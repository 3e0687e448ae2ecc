"""ctypes binding of libgi.so (the C ABI in include/gi_intersector.h).

`Intersector` mirrors the reference's KernelBVHController call pattern
(src/GPUIntersector/KernelBVHController.h:36-95): set both meshes, optionally
move them, run the intersection, read sorted pairs and exact segments.
Loading fails loudly when libgi.so has not been built; there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgi.so")


class GiConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("candidate_reserve", C.c_uint32), ("host_threads", C.c_int32),
                ("reserved", C.c_int32)]


class GiOptions(C.Structure):
    _fields_ = [("output_on_device", C.c_int32), ("want_segments", C.c_int32), ("rebuild", C.c_int32),
                ("use_graph", C.c_int32), ("stage_timing", C.c_int32), ("reserved", C.c_int32)]


_STAT_U64 = ["traversal_hits", "candidates", "filter_no", "filter_green", "filter_yellow", "exact_pairs",
             "exact_yellow_yes", "exact_yellow_no", "host_pairs", "host_yes", "pairs", "segments", "coplanar",
             "degenerate", "filter_predicates", "exact_predicates", "host_predicates"]
_STAT_F32 = ["ms_total", "ms_build", "ms_traverse", "ms_filter", "ms_exact", "ms_output", "ms_host",
             "ms_bvh_chunks", "ms_stage_meshes", "pad_ms"]

TIME_STAGING, TIME_BVH_CHUNKS, TIME_TRAVERSE, TIME_NONE = 1, 2, 4, 8


class GiStats(C.Structure):
    _fields_ = ([(n, C.c_uint64) for n in _STAT_U64] + [("grid", C.c_int32), ("retries", C.c_int32)]
                + [(n, C.c_float) for n in _STAT_F32] + [("bvh_chunk_launches", C.c_int32), ("bvh_chunk_leaves", C.c_uint32),
                                                              ("graph", C.c_int32), ("pad", C.c_int32)])

    def as_dict(self):
        d = {n: int(getattr(self, n)) for n in _STAT_U64}
        d.update({n: float(getattr(self, n)) for n in _STAT_F32})
        d["grid"] = int(self.grid)
        d["retries"] = int(self.retries)
        d["bvh_chunk_launches"] = int(self.bvh_chunk_launches)
        d["bvh_chunk_leaves"] = int(self.bvh_chunk_leaves)
        d["graph"] = int(self.graph)
        return d


class GiResult(C.Structure):
    _fields_ = [("n_pairs", C.c_uint64), ("pairs", C.POINTER(C.c_uint32)), ("seg_offsets", C.POINTER(C.c_uint32)),
                ("n_segments", C.c_uint64), ("exact_limbs", C.c_int32), ("on_device", C.c_int32),
                ("seg_exact", C.POINTER(C.c_uint64)), ("seg_exp2", C.POINTER(C.c_int32)),
                ("seg_f64", C.POINTER(C.c_double)), ("stats", GiStats), ("internal", C.c_void_p)]


class GiTuning(C.Structure):
    _fields_ = [("traverse_variant", C.c_int32), ("filter_blocks", C.c_int32), ("exact_blocks", C.c_int32),
                ("build_chunk", C.c_int32)]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class GiMeshPair(C.Structure):
    _fields_ = [("xyzA", C.POINTER(C.c_double)), ("triA", C.POINTER(C.c_uint32)), ("nvA", C.c_uint32),
                ("ntA", C.c_uint32), ("xyzB", C.POINTER(C.c_double)), ("triB", C.POINTER(C.c_uint32)),
                ("nvB", C.c_uint32), ("ntB", C.c_uint32)]


class GiOffMesh(C.Structure):
    _fields_ = [("soup_points", C.c_uint64), ("soup_polygons", C.c_uint64), ("extra_fan_triangles", C.c_uint64),
                ("n_dropped_degenerate", C.c_uint64), ("n_dropped_duplicate", C.c_uint64), ("nv", C.c_uint32),
                ("nt", C.c_uint32), ("xyz", C.POINTER(C.c_double)), ("tri", C.POINTER(C.c_uint32)),
                ("face_to_row", C.POINTER(C.c_uint32)), ("dropped_degenerate", C.POINTER(C.c_uint32)),
                ("dropped_duplicate", C.POINTER(C.c_uint32))]


class GiServerConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("slots", C.c_int32), ("snap_to_grid", C.c_int32),
                ("want_segments", C.c_int32), ("hbm_reserve_bytes", C.c_uint64)]


class GiRequestInfo(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("service_ms", C.c_double), ("latency_ms", C.c_double),
                ("t_admit_ms", C.c_double), ("t_done_ms", C.c_double), ("slot", C.c_int32), ("status", C.c_int32),
                ("refusals", C.c_uint32), ("pad", C.c_uint32)]


class GiError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"gi status {status}: {msg}")
        self.status = status


_lib = None

SYMBOLS = ["gi_create", "gi_destroy", "gi_last_error", "gi_status_string", "gi_set_mesh", "gi_set_mesh_device",
           "gi_set_transform", "gi_set_transform_matrix", "gi_build", "gi_get_vertices", "gi_get_tree", "gi_intersect",
           "gi_result_free",
           "gi_intersect_oneshot", "gi_intersect_noalloc", "gi_set_batch", "gi_intersect_batch", "gi_share_mesh", "gi_get_tuning",
           "gi_set_tuning", "gi_autotune", "gi_clear_device",
           "gi_reconstruct_device",
           "gi_server_create", "gi_server_add_mesh", "gi_server_submit", "gi_server_wait", "gi_server_destroy",
           "gi_exact_pair_host", "gi_off_read", "gi_off_free", "gi_off_write", "gi_pairs_csv_write",
           "gi_pairs_diff_csv_write", "gi_segments_csv_write", "gi_segments_obj_write", "gi_meshio_last_error"]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, dp, up = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint32)
        L.gi_create.argtypes = [C.POINTER(GiConfig), C.POINTER(vp)]
        L.gi_destroy.argtypes = [vp]
        L.gi_last_error.argtypes = [vp]
        L.gi_last_error.restype = C.c_char_p
        L.gi_status_string.argtypes = [C.c_int]
        L.gi_status_string.restype = C.c_char_p
        L.gi_set_mesh.argtypes = [vp, C.c_int, dp, C.c_uint32, up, C.c_uint32]
        L.gi_set_mesh_device.argtypes = [vp, C.c_int, vp, C.c_uint32, vp, C.c_uint32]
        L.gi_set_transform.argtypes = [vp, C.c_int, dp, dp, dp, C.c_int]
        L.gi_set_transform_matrix.argtypes = [vp, C.c_int, dp, C.c_int]
        L.gi_build.argtypes = [vp]
        L.gi_get_vertices.argtypes = [vp, C.c_int, dp, C.c_uint32]
        u32p = C.POINTER(C.c_uint32)
        L.gi_get_tree.argtypes = [vp, C.c_int, C.c_uint32, u32p, u32p, u32p, u32p, u32p, u32p]
        L.gi_intersect.argtypes = [vp, C.POINTER(GiOptions), C.POINTER(GiResult)]
        L.gi_result_free.argtypes = [C.POINTER(GiResult)]
        L.gi_intersect_oneshot.argtypes = [C.POINTER(GiConfig), dp, C.c_uint32, up, C.c_uint32, dp, C.c_uint32, up,
                                           C.c_uint32, C.POINTER(GiOptions), C.POINTER(GiResult)]
        u64p = C.POINTER(C.c_uint64)
        L.gi_intersect_noalloc.argtypes = [C.c_int32, dp, C.c_uint32, up, C.c_uint32, dp, C.c_uint32, up, C.c_uint32,
                                           C.c_int32, up, C.c_uint64, up, u64p, C.POINTER(C.c_int32), C.c_uint64,
                                           u64p, u64p, C.POINTER(C.c_int32), C.POINTER(GiStats)]
        L.gi_clear_device.argtypes = [vp]
        L.gi_reconstruct_device.argtypes = [vp]
        ip = C.POINTER(C.c_int32)
        L.gi_exact_pair_host.argtypes = [dp, dp, ip, ip, ip, ip, ip, C.POINTER(C.c_uint64), C.c_int32]
        L.gi_share_mesh.argtypes = [vp, C.c_int, vp, C.c_int]
        L.gi_get_tuning.argtypes = [vp, C.POINTER(GiTuning)]
        L.gi_set_tuning.argtypes = [vp, C.POINTER(GiTuning)]
        L.gi_autotune.argtypes = [vp, C.POINTER(GiOptions), C.c_int32, C.POINTER(GiTuning)]
        cp = C.c_char_p
        L.gi_off_read.argtypes = [cp, C.POINTER(GiOffMesh), C.POINTER(vp)]
        L.gi_off_free.argtypes = [vp]
        L.gi_off_write.argtypes = [cp, dp, C.c_uint32, up, C.c_uint32]
        L.gi_pairs_csv_write.argtypes = [cp, up, C.c_uint64, dp, C.c_uint32, up, C.c_uint32, dp, C.c_uint32, up,
                                         C.c_uint32]
        L.gi_pairs_diff_csv_write.argtypes = [cp, up, C.c_uint64, C.c_uint64, up, C.c_uint64, C.c_uint64, cp, dp,
                                              C.c_uint32, up, C.c_uint32, dp, C.c_uint32, up, C.c_uint32]
        L.gi_segments_csv_write.argtypes = [cp, C.POINTER(GiResult)]
        L.gi_segments_obj_write.argtypes = [cp, C.POINTER(GiResult)]
        L.gi_meshio_last_error.restype = C.c_char_p
        L.gi_set_batch.argtypes = [vp, C.POINTER(GiMeshPair), C.c_uint32]
        L.gi_intersect_batch.argtypes = [vp, C.POINTER(GiOptions), C.POINTER(GiResult), C.POINTER(C.c_uint64)]
        L.gi_server_create.argtypes = [C.POINTER(GiServerConfig), dp, C.c_uint32, up, C.c_uint32, C.POINTER(vp)]
        L.gi_server_add_mesh.argtypes = [vp, dp, C.c_uint32, up, C.c_uint32, ip]
        L.gi_server_submit.argtypes = [vp, C.c_int32, dp, C.POINTER(C.c_uint64)]
        L.gi_server_wait.argtypes = [vp, C.c_uint64, C.POINTER(GiResult), C.POINTER(GiRequestInfo)]
        L.gi_server_destroy.argtypes = [vp]
        for name in SYMBOLS:
            if name not in ("gi_last_error", "gi_status_string", "gi_meshio_last_error"):
                getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _limbs_to_int(words) -> int:
    v = 0
    for i, w in enumerate(words):
        v |= int(w) << (64 * i)
    n = 64 * len(words)
    if v >> (n - 1):
        v -= 1 << n
    return v


def decode_segments(exact: np.ndarray, exp2, offsets: np.ndarray):
    """exact: (S, 8, limbs) uint64. Returns per pair a sorted list of canonical
    ((x0,y0,z0),(x1,y1,z1)) Fraction endpoints."""
    out = []
    S = exact.shape[0]
    pts = []
    for s in range(S):
        nums = [_limbs_to_int(exact[s, e]) for e in range(8)]
        sc = Fraction(2) ** int(exp2[s]) if exp2 is not None else Fraction(1)
        p = tuple(Fraction(nums[k], nums[3]) * sc for k in range(3))
        q = tuple(Fraction(nums[4 + k], nums[7]) * sc for k in range(3))
        pts.append((p, q) if p <= q else (q, p))
    for i in range(len(offsets) - 1):
        out.append(sorted(pts[offsets[i]:offsets[i + 1]]))
    return out


def _check(st, ctx=None):
    if st != 0:
        msg = lib().gi_last_error(ctx).decode() if ctx else ""
        raise GiError(st, f"{lib().gi_status_string(st).decode()} {msg}")


class Intersector:
    """One gi_context: two mesh slots (A = 0, B = 1) on one GPU."""

    def __init__(self, device: int = 0, candidate_reserve: int = 0, host_threads: int = 0):
        self._ctx = C.c_void_p()
        cfg = GiConfig(device, candidate_reserve, host_threads, 0)
        _check(lib().gi_create(C.byref(cfg), C.byref(self._ctx)))

    def close(self):
        if self._ctx:
            lib().gi_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mesh(self, slot: int, xyz, tri):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        tri = np.ascontiguousarray(tri, dtype=np.uint32).reshape(-1, 3)
        _check(lib().gi_set_mesh(self._ctx, slot, xyz.ctypes.data_as(C.POINTER(C.c_double)), xyz.shape[0],
                                 tri.ctypes.data_as(C.POINTER(C.c_uint32)), tri.shape[0]), self._ctx)

    def set_mesh_device(self, slot: int, xyz_ptr: int, nv: int, tri_ptr: int, nt: int):
        _check(lib().gi_set_mesh_device(self._ctx, slot, C.c_void_p(xyz_ptr), nv, C.c_void_p(tri_ptr), nt),
               self._ctx)

    def set_transform(self, slot: int, rot_deg, trans, center, snap: bool = False):
        a = (C.c_double * 3)(*rot_deg)
        b = (C.c_double * 3)(*trans)
        c = (C.c_double * 3)(*center)
        _check(lib().gi_set_transform(self._ctx, slot, a, b, c, int(snap)), self._ctx)

    def set_transform_matrix(self, slot: int, m, snap: bool = False):
        arr = (C.c_double * 12)(*np.asarray(m, dtype=np.float64).ravel())
        _check(lib().gi_set_transform_matrix(self._ctx, slot, arr, int(snap)), self._ctx)

    def build(self):
        _check(lib().gi_build(self._ctx), self._ctx)

    def get_vertices(self, slot: int, nv: int) -> np.ndarray:
        """The slot's vertices as the predicates see them (moved, snapped)."""
        out = np.zeros((nv, 3), dtype=np.float64)
        _check(lib().gi_get_vertices(self._ctx, slot, out.ctypes.data_as(C.POINTER(C.c_double)), nv), self._ctx)
        return out

    def get_tree(self, slot: int, nt: int, tree: bool = True) -> dict:
        """The slot's sorted Morton keys and ids and, with tree, each internal
        node's left-child word, rope and leaf range and each leaf's rope."""
        out = {"keys": np.zeros(nt, np.uint32), "ids": np.zeros(nt, np.uint32)}
        if tree:
            m = max(nt - 1, 0)
            out.update(left=np.zeros(m, np.uint32), rope=np.zeros(m, np.uint32), range=np.zeros((m, 2), np.uint32),
                       leaf_rope=np.zeros(nt, np.uint32))
        ptr = lambda k: out[k].ctypes.data_as(C.POINTER(C.c_uint32)) if k in out else None
        _check(lib().gi_get_tree(self._ctx, slot, nt, ptr("keys"), ptr("ids"), ptr("left"), ptr("rope"),
                                 ptr("range"), ptr("leaf_rope")), self._ctx)
        return out

    def clear_device(self):
        _check(lib().gi_clear_device(self._ctx), self._ctx)

    def reconstruct_device(self):
        _check(lib().gi_reconstruct_device(self._ctx), self._ctx)

    def intersect_raw(self, output_on_device=False, want_segments=True, rebuild=False, use_graph=False,
                      stage_timing=0) -> GiResult:
        """Runs the query; the caller must gi_result_free the returned struct.
        stage_timing: 0 every stage timed; else a mask of TIME_STAGING,
        TIME_BVH_CHUNKS, TIME_TRAVERSE (TIME_NONE: none)."""
        res = GiResult()
        opt = GiOptions(int(output_on_device), int(want_segments), int(rebuild), int(use_graph), int(stage_timing), 0)
        _check(lib().gi_intersect(self._ctx, C.byref(opt), C.byref(res)), self._ctx)
        return res

    def intersect(self, want_segments=True, rebuild=False, decode=True, use_graph=False):
        res = self.intersect_raw(False, want_segments, rebuild, use_graph)
        try:
            return result_to_py(res, want_segments, decode)
        finally:
            lib().gi_result_free(C.byref(res))

    def get_tuning(self) -> dict:
        t = GiTuning()
        _check(lib().gi_get_tuning(self._ctx, C.byref(t)), self._ctx)
        return t.as_dict()

    def set_tuning(self, **kw):
        t = GiTuning()
        _check(lib().gi_get_tuning(self._ctx, C.byref(t)), self._ctx)
        for k, v in kw.items():
            setattr(t, k, int(v))
        _check(lib().gi_set_tuning(self._ctx, C.byref(t)), self._ctx)

    def autotune(self, reps=3, want_segments=True, rebuild=False) -> dict:
        t = GiTuning()
        opt = GiOptions(1, int(want_segments), int(rebuild), 0)
        _check(lib().gi_autotune(self._ctx, C.byref(opt), reps, C.byref(t)), self._ctx)
        return t.as_dict()

    def set_batch(self, pairs):
        """pairs: [((xyzA, triA), (xyzB, triB)), ...] host arrays."""
        arr = (GiMeshPair * len(pairs))()
        keep = []
        for i, ((xa, ta), (xb, tb)) in enumerate(pairs):
            xa = np.ascontiguousarray(xa, dtype=np.float64).reshape(-1, 3)
            ta = np.ascontiguousarray(ta, dtype=np.uint32).reshape(-1, 3)
            xb = np.ascontiguousarray(xb, dtype=np.float64).reshape(-1, 3)
            tb = np.ascontiguousarray(tb, dtype=np.uint32).reshape(-1, 3)
            keep += [xa, ta, xb, tb]
            arr[i] = GiMeshPair(xa.ctypes.data_as(C.POINTER(C.c_double)), ta.ctypes.data_as(C.POINTER(C.c_uint32)),
                                xa.shape[0], ta.shape[0], xb.ctypes.data_as(C.POINTER(C.c_double)),
                                tb.ctypes.data_as(C.POINTER(C.c_uint32)), xb.shape[0], tb.shape[0])
        _check(lib().gi_set_batch(self._ctx, arr, len(pairs)), self._ctx)
        self._n_batch = len(pairs)

    def intersect_batch_raw(self, output_on_device=False, want_segments=True, rebuild=False):
        """(GiResult, pair_offsets); the caller must gi_result_free the struct."""
        res = GiResult()
        offs = np.zeros(self._n_batch + 1, dtype=np.uint64)
        opt = GiOptions(int(output_on_device), int(want_segments), int(rebuild), 0)
        _check(lib().gi_intersect_batch(self._ctx, C.byref(opt), C.byref(res),
                                        offs.ctypes.data_as(C.POINTER(C.c_uint64))), self._ctx)
        return res, offs

    def intersect_batch(self, want_segments=True, rebuild=False, decode=True):
        """Per mesh pair results: a list of dicts like intersect()'s, plus the
        whole-batch stats."""
        res, offs = self.intersect_batch_raw(False, want_segments, rebuild)
        try:
            whole = result_to_py(res, want_segments, decode)
        finally:
            lib().gi_result_free(C.byref(res))
        out = []
        for i in range(self._n_batch):
            lo, hi = int(offs[i]), int(offs[i + 1])
            r = {"pairs": whole["pairs"][lo:hi]}
            if want_segments and decode:
                r["segments"] = whole["segments"][lo:hi]
            out.append(r)
        return out, whole["stats"]


def intersect_oneshot(xa, ta, xb, tb, want_segments=True, device=0, decode=True):
    """gi_intersect_oneshot: both meshes copied, built and intersected in one
    call (the reference's runGridForestIntersectionPipeline)."""
    xa = np.ascontiguousarray(xa, dtype=np.float64).reshape(-1, 3)
    ta = np.ascontiguousarray(ta, dtype=np.uint32).reshape(-1, 3)
    xb = np.ascontiguousarray(xb, dtype=np.float64).reshape(-1, 3)
    tb = np.ascontiguousarray(tb, dtype=np.uint32).reshape(-1, 3)
    cfg = GiConfig(device, 0, 0, 0)
    opt = GiOptions(0, int(want_segments), 0, 0)
    res = GiResult()
    dp, up = C.POINTER(C.c_double), C.POINTER(C.c_uint32)
    _check(lib().gi_intersect_oneshot(C.byref(cfg), xa.ctypes.data_as(dp), xa.shape[0], ta.ctypes.data_as(up),
                                      ta.shape[0], xb.ctypes.data_as(dp), xb.shape[0], tb.ctypes.data_as(up),
                                      tb.shape[0], C.byref(opt), C.byref(res)))
    try:
        return result_to_py(res, want_segments, decode)
    finally:
        lib().gi_result_free(C.byref(res))


def intersect_noalloc(xa, ta, xb, tb, want_segments=True, device=0, pairs_cap=None, seg_cap=None):
    """gi_intersect_noalloc into buffers this wrapper allocates (sized for
    every triangle pair of the smaller mesh times 6 segments unless given).
    Returns (status, result dict or None, needed (pairs, segments))."""
    xa = np.ascontiguousarray(xa, dtype=np.float64).reshape(-1, 3)
    ta = np.ascontiguousarray(ta, dtype=np.uint32).reshape(-1, 3)
    xb = np.ascontiguousarray(xb, dtype=np.float64).reshape(-1, 3)
    tb = np.ascontiguousarray(tb, dtype=np.uint32).reshape(-1, 3)
    if pairs_cap is None:
        pairs_cap = 8 * (ta.shape[0] + tb.shape[0]) + 1024
    if seg_cap is None:
        seg_cap = pairs_cap
    pairs = np.zeros(2 * max(pairs_cap, 1), dtype=np.uint32)
    offs = np.zeros(pairs_cap + 1, dtype=np.uint32)
    exact = np.zeros(max(seg_cap, 1) * 8 * 4, dtype=np.uint64)
    exp2 = np.zeros(max(seg_cap, 1), dtype=np.int32)
    n_pairs, n_seg, limbs = C.c_uint64(), C.c_uint64(), C.c_int32()
    st = GiStats()
    dp, up = C.POINTER(C.c_double), C.POINTER(C.c_uint32)
    rc = lib().gi_intersect_noalloc(device, xa.ctypes.data_as(dp), xa.shape[0], ta.ctypes.data_as(up), ta.shape[0],
                                    xb.ctypes.data_as(dp), xb.shape[0], tb.ctypes.data_as(up), tb.shape[0],
                                    int(want_segments), pairs.ctypes.data_as(up), pairs_cap, offs.ctypes.data_as(up),
                                    exact.ctypes.data_as(C.POINTER(C.c_uint64)),
                                    exp2.ctypes.data_as(C.POINTER(C.c_int32)), seg_cap, C.byref(n_pairs),
                                    C.byref(n_seg), C.byref(limbs), C.byref(st))
    need = (int(n_pairs.value), int(n_seg.value))
    if rc != 0:
        return rc, None, need
    n, S, L = need[0], need[1], int(limbs.value)
    out = dict(pairs=pairs[: 2 * n].reshape(n, 2).copy(), seg_offsets=offs[: n + 1].copy(), stats=st.as_dict(),
               n_segments=S, exact_limbs=L)
    if want_segments:
        ex = exact[: S * 8 * L].reshape(S, 8, L)
        out["seg_exact"] = ex.copy()
        out["seg_exp2"] = exp2[:S].copy()
        out["segments"] = decode_segments(ex, exp2[:S], out["seg_offsets"]) if S else [[] for _ in range(n)]
    return rc, out, need


def result_to_py(res: GiResult, want_segments=True, decode=True):
    n = int(res.n_pairs)
    S = int(res.n_segments)
    pairs = np.ctypeslib.as_array(res.pairs, shape=(max(2 * n, 1),))[: 2 * n].reshape(n, 2).copy() if n else \
        np.zeros((0, 2), dtype=np.uint32)
    offs = np.ctypeslib.as_array(res.seg_offsets, shape=(n + 1,)).copy()
    out = dict(pairs=pairs, seg_offsets=offs, stats=res.stats.as_dict(), n_segments=S)
    if want_segments and S:
        L = int(res.exact_limbs)
        exact = np.ctypeslib.as_array(res.seg_exact, shape=(S * 8 * L,)).reshape(S, 8, L).copy()
        exp2 = np.ctypeslib.as_array(res.seg_exp2, shape=(S,)).copy() if res.seg_exp2 else None
        f64 = np.ctypeslib.as_array(res.seg_f64, shape=(S * 6,)).reshape(S, 6).copy()
        out.update(seg_exact=exact, seg_exp2=exp2, seg_f64=f64)
        if decode:
            out["segments"] = decode_segments(exact, exp2, offs)
    elif want_segments:
        out["segments"] = [[] for _ in range(n)]
    return out


def _io_check(st):
    if st != 0:
        raise GiError(st, f"{lib().gi_status_string(st).decode()}: {lib().gi_meshio_last_error().decode()}")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _mesh_arrays(xyz, tri):
    return (np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3),
            np.ascontiguousarray(tri, dtype=np.uint32).reshape(-1, 3))


def read_off(path):
    """OFF -> dict(xyz, tri, face_to_row, dropped_degenerate, dropped_duplicate,
    soup_points, soup_polygons, extra_fan_triangles), with the canonical face
    order of the reference's loader."""
    m = GiOffMesh()
    owner = C.c_void_p()
    _io_check(lib().gi_off_read(os.fsencode(path), C.byref(m), C.byref(owner)))
    try:
        def arr(p, n, shape):
            if n == 0:
                return np.zeros(shape, dtype=np.float64 if p._type_ is C.c_double else np.uint32)
            return np.ctypeslib.as_array(p, shape=(n,)).copy().reshape(shape)
        return dict(xyz=arr(m.xyz, 3 * m.nv, (m.nv, 3)), tri=arr(m.tri, 3 * m.nt, (m.nt, 3)),
                    face_to_row=arr(m.face_to_row, m.nt, (m.nt,)),
                    dropped_degenerate=arr(m.dropped_degenerate, m.n_dropped_degenerate, (m.n_dropped_degenerate,)),
                    dropped_duplicate=arr(m.dropped_duplicate, m.n_dropped_duplicate, (m.n_dropped_duplicate,)),
                    soup_points=m.soup_points, soup_polygons=m.soup_polygons,
                    extra_fan_triangles=m.extra_fan_triangles)
    finally:
        lib().gi_off_free(owner)


def write_off(path, xyz, tri):
    xyz, tri = _mesh_arrays(xyz, tri)
    _io_check(lib().gi_off_write(os.fsencode(path), _dp(xyz), xyz.shape[0], _up(tri), tri.shape[0]))


def write_pairs_csv(path, pairs, mesh_a, mesh_b):
    """The reference's run CSV (FaceA, FaceB, vertex indices, coordinates)."""
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    (xa, ta), (xb, tb) = _mesh_arrays(*mesh_a), _mesh_arrays(*mesh_b)
    _io_check(lib().gi_pairs_csv_write(os.fsencode(path), _up(pairs), pairs.shape[0], _dp(xa), xa.shape[0], _up(ta),
                                       ta.shape[0], _dp(xb), xb.shape[0], _up(tb), tb.shape[0]))


def write_pairs_diff_csv(path, pairs1, id1, pairs2, id2, filt, mesh_a, mesh_b):
    p1 = np.ascontiguousarray(pairs1, dtype=np.uint32).reshape(-1, 2)
    p2 = np.ascontiguousarray(pairs2, dtype=np.uint32).reshape(-1, 2)
    (xa, ta), (xb, tb) = _mesh_arrays(*mesh_a), _mesh_arrays(*mesh_b)
    _io_check(lib().gi_pairs_diff_csv_write(os.fsencode(path), _up(p1), p1.shape[0], id1, _up(p2), p2.shape[0], id2,
                                            filt.encode(), _dp(xa), xa.shape[0], _up(ta), ta.shape[0], _dp(xb),
                                            xb.shape[0], _up(tb), tb.shape[0]))


def write_segments(res: GiResult, path, fmt="csv"):
    """Exact segments of a host result: 'csv' (exact integers + doubles) or 'obj' (polyline)."""
    fn = lib().gi_segments_csv_write if fmt == "csv" else lib().gi_segments_obj_write
    _io_check(fn(os.fsencode(path), C.byref(res)))


class Server:
    """gi_server: one resident static mesh, registered moving meshes, queued
    rigid-transform requests processed by `slots` concurrent GPU slots."""

    def __init__(self, xyz, tri, slots=4, snap=True, want_segments=True, device=0, hbm_reserve_bytes=0):
        self._s = C.c_void_p()
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        tri = np.ascontiguousarray(tri, dtype=np.uint32).reshape(-1, 3)
        cfg = GiServerConfig(device, slots, int(snap), int(want_segments), hbm_reserve_bytes)
        _check(lib().gi_server_create(C.byref(cfg), xyz.ctypes.data_as(C.POINTER(C.c_double)), xyz.shape[0],
                                      tri.ctypes.data_as(C.POINTER(C.c_uint32)), tri.shape[0], C.byref(self._s)))

    def add_mesh(self, xyz, tri) -> int:
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        tri = np.ascontiguousarray(tri, dtype=np.uint32).reshape(-1, 3)
        mid = C.c_int32()
        _check(lib().gi_server_add_mesh(self._s, xyz.ctypes.data_as(C.POINTER(C.c_double)), xyz.shape[0],
                                        tri.ctypes.data_as(C.POINTER(C.c_uint32)), tri.shape[0], C.byref(mid)))
        return mid.value

    def submit(self, mesh_id: int, m) -> int:
        arr = (C.c_double * 12)(*np.asarray(m, dtype=np.float64).ravel())
        t = C.c_uint64()
        _check(lib().gi_server_submit(self._s, mesh_id, arr, C.byref(t)))
        return t.value

    def wait(self, ticket: int, decode=True):
        res = GiResult()
        info = GiRequestInfo()
        st = lib().gi_server_wait(self._s, ticket, C.byref(res), C.byref(info))
        try:
            _check(st)
            out = result_to_py(res, True, decode)
        finally:
            lib().gi_result_free(C.byref(res))
        out["info"] = {"queue_ms": info.queue_ms, "service_ms": info.service_ms, "latency_ms": info.latency_ms,
                       "t_admit_ms": info.t_admit_ms, "t_done_ms": info.t_done_ms, "slot": info.slot,
                       "refusals": info.refusals}
        return out

    def close(self):
        if self._s:
            lib().gi_server_destroy(self._s)
            self._s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def exact_pair_host(a, b, limbs_cap=64):
    """Host residual stage on one pair: (hit, coplanar, canonical segments)."""
    hit, cop, nseg, exp2, limbs = (C.c_int32() for _ in range(5))
    buf = (C.c_uint64 * (6 * 8 * limbs_cap))()
    da = (C.c_double * 9)(*[float(x) for x in a])
    db = (C.c_double * 9)(*[float(x) for x in b])
    _check(lib().gi_exact_pair_host(da, db, C.byref(hit), C.byref(cop), C.byref(nseg), C.byref(exp2), C.byref(limbs),
                                    buf, limbs_cap))
    raw = np.ctypeslib.as_array(buf).reshape(6, 8, limbs_cap)
    segs = decode_segments(raw[: nseg.value], [exp2.value] * nseg.value, np.array([0, nseg.value]))[0] \
        if nseg.value else []
    return bool(hit.value), bool(cop.value), segs
