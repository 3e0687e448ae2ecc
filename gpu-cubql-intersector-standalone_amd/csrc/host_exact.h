// host_exact.h -- host residual stage: exact intersection of triangle pairs
// whose coordinates are general doubles (not on the integer grid the GPU
// exact stage handles). Replaces the reference's CGAL fallback on TBB
// (src/CPU/YellowFilter.h:80-137) and CGAL segment construction
// (src/CPU/MeshIntersectionCurves.cpp:27-92).
#pragma once

#include <stdint.h>

#include <vector>

namespace gi {

struct HostHit {
    uint32_t a, b;
    int32_t coplanar, degenerate;
    int32_t nseg, exp2, limbs;
    std::vector<uint64_t> data;  // nseg * 8 * limbs
};

// Exact decision + segments for one pair (9 doubles each). Every pair of
// finite doubles fits (a pair first runs on 4,096-bit integers, then, if its
// exponent spread needs more, on 16,384-bit ones); *ok = false only for a
// non-finite coordinate. *npred (may be null) receives the orient3d and
// orient2d signs evaluated.
bool host_exact_pair(const double* a, const double* b, HostHit& out, bool* ok, int* npred = nullptr);

// host_exact_batch return codes
constexpr int kHostUnsound = -1;     // a pair the filter certified GREEN does not intersect
constexpr int kHostNonFinite = -2;   // a coordinate is infinite or NaN

// Runs host_exact_pair over a list of (a, b | status << 30) pairs with
// nthreads threads; keeps the intersecting ones, sorted by (a, b).
// vA/vB are xyz0-padded vertex arrays (4 doubles per vertex). predicates
// (may be null) receives the signs evaluated over the whole list.
int host_exact_batch(const double* vA, const uint32_t* triA, const double* vB, const uint32_t* triB,
                     const uint32_t* list, uint32_t n, int nthreads, std::vector<HostHit>& hits, uint64_t* yellow_yes,
                     uint64_t* predicates = nullptr);

}  // namespace gi
