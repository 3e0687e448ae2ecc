// filter.h -- FP64 certified filter for the triangle/triangle decision.
//
// Stage 1 of the cascade. It runs the reference's NO/GREEN/YELLOW decision
// tree (edgeTriWithSigns / classifyLocal,
// src/GPU_exact_predicates/OriginShiftedFloatPredicateCore.cuh:70-124), but
// every orient3d sign is certified by an FP64 midpoint-radius enclosure
// instead of FP32 intervals with CUDA's directed-rounding intrinsics
// (GPUPredicatesCommon.h:38-100), which HIP does not provide:
//
//   det computed in round-to-nearest, radius = (7 + 56 eps) eps * permanent,
//   eps = 2^-53   (J. R. Shewchuk, "Adaptive Precision Floating-Point
//   Arithmetic and Fast Robust Geometric Predicates", DCG 18, 1997, bound
//   errboundA of orient3d)
//
// The enclosure [det - radius, det + radius] contains the exact determinant
// of the double inputs, so a sign is certified when 0 lies outside it. A sign
// that cannot be certified (including an exact zero) is reported as 0 and
// makes the pair YELLOW, i.e. it goes to the exact integer stage.
//
// The code is compiled without FP contraction (no FMA): the bound assumes
// separately rounded products and sums.
#pragma once

#include <hip/hip_runtime.h>

namespace gi {

enum PairStatus { PAIR_NO = 0, PAIR_GREEN = 1, PAIR_YELLOW = 2 };

struct d3 { double x, y, z; };

__device__ __forceinline__ d3 d3sub(const d3& a, const d3& b) {
    return d3{__dsub_rn(a.x, b.x), __dsub_rn(a.y, b.y), __dsub_rn(a.z, b.z)};
}

// Certified sign of (a-d).((b-d)x(c-d)); 0 when not certified.
__device__ __forceinline__ int orient3d_cert(const d3& a, const d3& b, const d3& c, const d3& d) {
    const double kBound = 7.7715611723761027e-16 + 6.9e-31;  // (7 + 56 eps) eps, rounded up
    d3 ad = d3sub(a, d), bd = d3sub(b, d), cd = d3sub(c, d);
    double bdxcdy = __dmul_rn(bd.x, cd.y), cdxbdy = __dmul_rn(cd.x, bd.y);
    double cdxady = __dmul_rn(cd.x, ad.y), adxcdy = __dmul_rn(ad.x, cd.y);
    double adxbdy = __dmul_rn(ad.x, bd.y), bdxady = __dmul_rn(bd.x, ad.y);
    double det = __dadd_rn(__dadd_rn(__dmul_rn(ad.z, __dsub_rn(bdxcdy, cdxbdy)),
                                     __dmul_rn(bd.z, __dsub_rn(cdxady, adxcdy))),
                           __dmul_rn(cd.z, __dsub_rn(adxbdy, bdxady)));
    double perm = __dadd_rn(__dadd_rn(__dmul_rn(__dadd_rn(fabs(bdxcdy), fabs(cdxbdy)), fabs(ad.z)),
                                      __dmul_rn(__dadd_rn(fabs(cdxady), fabs(adxcdy)), fabs(bd.z))),
                            __dmul_rn(__dadd_rn(fabs(adxbdy), fabs(bdxady)), fabs(cd.z)));
    // below ~1e-270 products may have underflowed and the bound is void
    if (!(perm >= 1e-270) || !(perm <= 1e300)) return 0;
    double err = __dmul_rn(kBound, perm);
    if (det > err) return 1;
    if (-det > err) return -1;
    return 0;
}

// edgeTriWithSigns (OriginShiftedFloatPredicateCore.cuh:70-85)
__device__ __forceinline__ int edge_tri_cert(const d3& a, const d3& b, int sa, int sb, const d3& p, const d3& q,
                                             const d3& r, int& np) {
    if (sa == 0 || sb == 0) return PAIR_YELLOW;
    if (sa == sb) return PAIR_NO;
    np += 3;
    int e0 = orient3d_cert(a, b, p, q);
    int e1 = orient3d_cert(a, b, q, r);
    int e2 = orient3d_cert(a, b, r, p);
    if (e0 == 0 || e1 == 0 || e2 == 0) return PAIR_YELLOW;
    if ((e0 > 0 && e1 > 0 && e2 > 0) || (e0 < 0 && e1 < 0 && e2 < 0)) return PAIR_GREEN;
    return PAIR_NO;
}

// classifyLocal (OriginShiftedFloatPredicateCore.cuh:91-124), absolute
// coordinates: the certified orient3d needs no anchor shift. np counts the
// orient3d evaluations.
__device__ __forceinline__ int classify_pair_fp64(const d3* A, const d3* B, int& np) {
    int sB0 = orient3d_cert(A[0], A[1], A[2], B[0]);
    int sB1 = orient3d_cert(A[0], A[1], A[2], B[1]);
    int sB2 = orient3d_cert(A[0], A[1], A[2], B[2]);
    np += 3;
    if (sB0 != 0 && sB0 == sB1 && sB1 == sB2) return PAIR_NO;
    int sA0 = orient3d_cert(B[0], B[1], B[2], A[0]);
    int sA1 = orient3d_cert(B[0], B[1], B[2], A[1]);
    int sA2 = orient3d_cert(B[0], B[1], B[2], A[2]);
    np += 3;
    if (sA0 != 0 && sA0 == sA1 && sA1 == sA2) return PAIR_NO;
    int r;
    r = edge_tri_cert(A[0], A[1], sA0, sA1, B[0], B[1], B[2], np); if (r != PAIR_NO) return r;
    r = edge_tri_cert(A[1], A[2], sA1, sA2, B[0], B[1], B[2], np); if (r != PAIR_NO) return r;
    r = edge_tri_cert(A[2], A[0], sA2, sA0, B[0], B[1], B[2], np); if (r != PAIR_NO) return r;
    r = edge_tri_cert(B[0], B[1], sB0, sB1, A[0], A[1], A[2], np); if (r != PAIR_NO) return r;
    r = edge_tri_cert(B[1], B[2], sB1, sB2, A[0], A[1], A[2], np); if (r != PAIR_NO) return r;
    r = edge_tri_cert(B[2], B[0], sB2, sB0, A[0], A[1], A[2], np); if (r != PAIR_NO) return r;
    return PAIR_NO;
}

}  // namespace gi
