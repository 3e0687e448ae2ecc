/*
 * oracle.h -- CPU oracle for the exact triangle-mesh intersection path.
 *
 * TEST INFRASTRUCTURE ONLY: this library is the checker that tests/, the
 * smoke check and the CPU-baseline leg of bench.py compare the HIP product
 * against. The product never links, loads or calls it.
 *
 * What it restates (reference files, relative to the reference root):
 *   - candidate boxes: exact closed double boxes of each triangle, the boxes
 *     CGAL's box_intersection_d sees in the reference CPU path
 *     (src/CPU/TestSuite.cpp:79-102, apps/mainApplication.cpp:1187-1189);
 *     the GPU path pads FP32 boxes (src/common/utils.cu:595-643) only to get a
 *     superset of these;
 *   - floating filter: the NO/GREEN/YELLOW decision tree of
 *     src/GPU_exact_predicates/OriginShiftedFloatPredicateCore.cuh:70-124
 *     (edgeTriWithSigns, classifyLocal), evaluated with a certified FP64
 *     orient3d (Shewchuk's semi-static bound) instead of FP32 intervals;
 *   - exact decision: classifyPairBigIntV2 and helpers
 *     (src/GPU_exact_predicates/ExactBigIntPredicate.cu:217-489) on exact
 *     integers, with no exponent-spread cap and with degenerate triangles
 *     decided by point-set semantics (the reference hands them to CGAL);
 *   - segments: the shapes CGAL::intersection(Triangle_3, Triangle_3) returns
 *     in src/CPU/MeshIntersectionCurves.cpp:63-77, computed exactly: one
 *     segment (or a point) for non-coplanar pairs, the boundary edges of the
 *     overlap polygon for coplanar pairs.
 */
#ifndef GI_ORACLE_H
#define GI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_SEG 6

/* One exact homogeneous endpoint set is exported per segment as 8 numbers
 * (x0,y0,z0,w0,x1,y1,z1,w1), each `limbs` little-endian two's complement
 * 64-bit limbs. The true coordinate is (x/w) * 2^exp2. */
typedef struct {
    int32_t intersect; /* 0 or 1 */
    int32_t coplanar;  /* 1 if both triangles are non-degenerate and coplanar */
    int32_t degenerate;/* 1 if either triangle has zero area */
    int32_t nseg;      /* 0..ORACLE_MAX_SEG */
    int32_t exp2;      /* common power-of-two scale of the exported numbers */
    int32_t limbs;     /* limbs per exported number (<= ORACLE_EXPORT_LIMBS) */
    int32_t filter_status; /* -1 not run, 0 NO, 1 GREEN, 2 YELLOW */
    int32_t pad;
} oracle_tt_info;

#define ORACLE_EXPORT_LIMBS 24

typedef struct {
    oracle_tt_info info;
    uint64_t seg[ORACLE_MAX_SEG][8][ORACLE_EXPORT_LIMBS];
} oracle_tt_result;

/* Decide (and, if they meet, compute the exact intersection of) two closed
 * triangles given as 9 doubles each. use_filter selects the FP64 filter fast
 * path before the exact stage. Returns 0 on success. */
int oracle_tritri(const double a[9], const double b[9], int use_filter, oracle_tt_result* out);

/* Sign of the exact orient3d determinant (a-d).((b-d)x(c-d)). */
int oracle_orient3d_sign(const double a[3], const double b[3], const double c[3], const double d[3]);
/* FP64 semi-static orient3d: returns +1/-1 when certified, 0 when uncertain. */
int oracle_orient3d_filter(const double a[3], const double b[3], const double c[3], const double d[3]);

typedef struct {
    const double* xyz;   /* 3 * nv */
    uint32_t nv;
    const uint32_t* tri; /* 3 * nt */
    uint32_t nt;
} oracle_mesh;

/* Candidate pairs = closed exact double-box overlaps. method 0: sort and
 * sweep, 1: brute force. Writes up to cap pairs (a,b) into out (may be NULL)
 * in unspecified order; returns the total count. */
int64_t oracle_candidates(const oracle_mesh* A, const oracle_mesh* B, int method, int nthreads,
                          uint32_t* out, int64_t cap);

typedef struct {
    int64_t n_candidates;
    int64_t n_filter_no, n_filter_green, n_filter_yellow;
    int64_t n_exact;          /* pairs that ran the exact stage */
    int64_t n_pairs;          /* intersecting pairs */
    int64_t n_coplanar;       /* intersecting coplanar pairs */
    int64_t n_degenerate;     /* intersecting pairs with a zero-area triangle */
    int64_t n_segments;
    uint32_t* pairs;          /* 2 * n_pairs, sorted by (a, b) */
    uint32_t* seg_offsets;    /* n_pairs + 1 */
    int32_t* seg_exp2;        /* per pair */
    int32_t limbs;            /* limbs per exported number */
    uint64_t* seg_data;       /* n_segments * 8 * limbs */
} oracle_result;

/* Full CPU path: candidates (sweep) -> filter -> exact -> sorted pairs (+ exact
 * segments when want_segments). a_limit restricts the A triangles considered
 * to [0, a_limit) (0 = all), which is how bounded CPU samples are taken. */
int oracle_intersect(const oracle_mesh* A, const oracle_mesh* B, int nthreads, int use_filter,
                     int want_segments, uint32_t a_limit, oracle_result* out);
void oracle_result_free(oracle_result* r);

#ifdef __cplusplus
}
#endif

#endif
