"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke check and by the
CPU-baseline leg of bench.py as the checker of the HIP product path. The
product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

MAX_SEG = 6
EXPORT_LIMBS = 24


class TTInfo(C.Structure):
    _fields_ = [
        ("intersect", C.c_int32),
        ("coplanar", C.c_int32),
        ("degenerate", C.c_int32),
        ("nseg", C.c_int32),
        ("exp2", C.c_int32),
        ("limbs", C.c_int32),
        ("filter_status", C.c_int32),
        ("pad", C.c_int32),
    ]


class TTResult(C.Structure):
    _fields_ = [("info", TTInfo), ("seg", C.c_uint64 * (MAX_SEG * 8 * EXPORT_LIMBS))]


class Mesh(C.Structure):
    _fields_ = [
        ("xyz", C.POINTER(C.c_double)),
        ("nv", C.c_uint32),
        ("tri", C.POINTER(C.c_uint32)),
        ("nt", C.c_uint32),
    ]


class Result(C.Structure):
    _fields_ = [
        ("n_candidates", C.c_int64),
        ("n_filter_no", C.c_int64),
        ("n_filter_green", C.c_int64),
        ("n_filter_yellow", C.c_int64),
        ("n_exact", C.c_int64),
        ("n_pairs", C.c_int64),
        ("n_coplanar", C.c_int64),
        ("n_degenerate", C.c_int64),
        ("n_segments", C.c_int64),
        ("pairs", C.POINTER(C.c_uint32)),
        ("seg_offsets", C.POINTER(C.c_uint32)),
        ("seg_exp2", C.POINTER(C.c_int32)),
        ("limbs", C.c_int32),
        ("seg_data", C.POINTER(C.c_uint64)),
    ]


_lib = None


def build() -> None:
    """Compile liboracle.so with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.oracle_tritri.argtypes = [dp, dp, C.c_int, C.POINTER(TTResult)]
        L.oracle_tritri.restype = C.c_int
        L.oracle_orient3d_sign.argtypes = [dp, dp, dp, dp]
        L.oracle_orient3d_sign.restype = C.c_int
        L.oracle_orient3d_filter.argtypes = [dp, dp, dp, dp]
        L.oracle_orient3d_filter.restype = C.c_int
        L.oracle_candidates.argtypes = [C.POINTER(Mesh), C.POINTER(Mesh), C.c_int, C.c_int,
                                        C.POINTER(C.c_uint32), C.c_int64]
        L.oracle_candidates.restype = C.c_int64
        L.oracle_intersect.argtypes = [C.POINTER(Mesh), C.POINTER(Mesh), C.c_int, C.c_int, C.c_int,
                                       C.c_uint32, C.POINTER(Result)]
        L.oracle_intersect.restype = C.c_int
        L.oracle_result_free.argtypes = [C.POINTER(Result)]
        L.oracle_result_free.restype = None
        _lib = L
    return _lib


def _dbl(seq):
    a = (C.c_double * len(seq))(*[float(x) for x in seq])
    return a


def _limbs_to_int(words) -> int:
    v = 0
    for i, w in enumerate(words):
        v |= int(w) << (64 * i)
    n = 64 * len(words)
    if v >> (n - 1):
        v -= 1 << n
    return v


def _point(nums, exp2):
    x, y, z, w = nums
    s = Fraction(2) ** exp2
    return (Fraction(x, w) * s, Fraction(y, w) * s, Fraction(z, w) * s)


def canon_segment(p, q):
    return (p, q) if p <= q else (q, p)


def tritri(a, b, use_filter=True):
    """Exact decision and intersection of two triangles (9 doubles each).

    Returns dict(intersect, coplanar, degenerate, filter_status, segments)
    with segments as a sorted list of canonical (P, Q) Fraction triples."""
    r = TTResult()
    rc = lib().oracle_tritri(_dbl(a), _dbl(b), 1 if use_filter else 0, C.byref(r))
    if rc != 0:
        raise RuntimeError(f"oracle_tritri failed: {rc}")
    info = r.info
    segs = []
    L = info.limbs
    raw = np.ctypeslib.as_array(r.seg).reshape(MAX_SEG, 8, EXPORT_LIMBS)
    for s in range(info.nseg):
        nums = [_limbs_to_int(raw[s, e, :L]) for e in range(8)]
        p = _point(nums[0:4], info.exp2)
        q = _point(nums[4:8], info.exp2)
        segs.append(canon_segment(p, q))
    return dict(intersect=bool(info.intersect), coplanar=bool(info.coplanar),
                degenerate=bool(info.degenerate), filter_status=info.filter_status,
                segments=sorted(segs))


def orient3d_sign(a, b, c, d) -> int:
    return lib().oracle_orient3d_sign(_dbl(a), _dbl(b), _dbl(c), _dbl(d))


def orient3d_filter(a, b, c, d) -> int:
    return lib().oracle_orient3d_filter(_dbl(a), _dbl(b), _dbl(c), _dbl(d))


def _mesh(xyz, tri):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    tri = np.ascontiguousarray(tri, dtype=np.uint32).reshape(-1, 3)
    m = Mesh(xyz.ctypes.data_as(C.POINTER(C.c_double)), xyz.shape[0],
             tri.ctypes.data_as(C.POINTER(C.c_uint32)), tri.shape[0])
    return m, (xyz, tri)


def candidates(xyzA, triA, xyzB, triB, method=0, nthreads=8):
    """All (a, b) with overlapping closed double boxes, sorted."""
    mA, keepA = _mesh(xyzA, triA)
    mB, keepB = _mesh(xyzB, triB)
    n = lib().oracle_candidates(C.byref(mA), C.byref(mB), method, nthreads, None, 0)
    out = np.zeros((max(n, 1), 2), dtype=np.uint32)
    lib().oracle_candidates(C.byref(mA), C.byref(mB), method, nthreads,
                            out.ctypes.data_as(C.POINTER(C.c_uint32)), n)
    out = out[:n]
    key = out[:, 0].astype(np.uint64) << np.uint64(32) | out[:, 1].astype(np.uint64)
    return out[np.argsort(key, kind="stable")]


def intersect(xyzA, triA, xyzB, triB, nthreads=8, use_filter=True, want_segments=True, a_limit=0,
              decode_segments=True):
    """Whole CPU path. Returns dict with sorted pairs (n,2 uint32), stats and,
    when decode_segments, a list (per pair) of sorted canonical segments."""
    mA, keepA = _mesh(xyzA, triA)
    mB, keepB = _mesh(xyzB, triB)
    r = Result()
    rc = lib().oracle_intersect(C.byref(mA), C.byref(mB), nthreads, 1 if use_filter else 0,
                                1 if want_segments else 0, a_limit, C.byref(r))
    try:
        if rc != 0:
            raise RuntimeError(f"oracle_intersect failed: {rc}")
        n = r.n_pairs
        pairs = np.ctypeslib.as_array(r.pairs, shape=(max(n, 1) * 2,))[: 2 * n].reshape(n, 2).copy()
        stats = {k: getattr(r, k) for k in ("n_candidates", "n_filter_no", "n_filter_green",
                                             "n_filter_yellow", "n_exact", "n_pairs", "n_coplanar",
                                             "n_degenerate", "n_segments")}
        out = dict(pairs=pairs, stats=stats)
        if want_segments:
            offs = np.ctypeslib.as_array(r.seg_offsets, shape=(n + 1,)).copy()
            out["seg_offsets"] = offs
            if decode_segments:
                L = r.limbs
                ns = r.n_segments
                data = np.ctypeslib.as_array(r.seg_data, shape=(max(ns, 1) * 8 * L,))[: ns * 8 * L]
                data = data.reshape(ns, 8, L)
                exps = np.ctypeslib.as_array(r.seg_exp2, shape=(max(n, 1),))[:n]
                segs = []
                for i in range(n):
                    lst = []
                    for s in range(offs[i], offs[i + 1]):
                        nums = [_limbs_to_int(data[s, e]) for e in range(8)]
                        lst.append(canon_segment(_point(nums[0:4], int(exps[i])),
                                                 _point(nums[4:8], int(exps[i]))))
                    segs.append(sorted(lst))
                out["segments"] = segs
        return out
    finally:
        lib().oracle_result_free(C.byref(r))
