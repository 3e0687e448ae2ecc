// KernelBVHController.cpp -- the C++ facade over the C ABI (see the header).
#include "KernelBVHController.h"

#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <utility>

#include "../../include/gi_intersector.h"

namespace gi_host {

namespace {

void check(int st, gi_context* ctx) {
    if (st == GI_OK) return;
    std::string msg = gi_status_string(st);
    if (ctx) {
        msg += ": ";
        msg += gi_last_error(ctx);
    }
    throw Error(st, msg);
}

Point3 centroid(const double3* v, int n) {
    Point3 c{0.0, 0.0, 0.0};
    if (n <= 0) return c;
    for (int i = 0; i < n; ++i) { c.x += v[i].x; c.y += v[i].y; c.z += v[i].z; }
    c.x /= n; c.y /= n; c.z /= n;
    return c;
}

// two's complement little-endian limbs -> decimal
std::string to_decimal(const uint64_t* limbs, int n) {
    bool neg = (limbs[n - 1] >> 63) != 0;
    std::vector<uint32_t> mag(2 * (size_t)n);
    uint64_t carry = 1;
    for (int i = 0; i < n; ++i) {
        uint64_t v = limbs[i];
        if (neg) {
            v = ~v;
            unsigned __int128 s = (unsigned __int128)v + carry;
            v = (uint64_t)s;
            carry = (uint64_t)(s >> 64);
        }
        mag[2 * i] = (uint32_t)v;
        mag[2 * i + 1] = (uint32_t)(v >> 32);
    }
    std::string digits;
    while (!mag.empty() && mag.back() == 0) mag.pop_back();
    if (mag.empty()) return "0";
    while (!mag.empty()) {
        uint64_t rem = 0;
        for (size_t i = mag.size(); i-- > 0;) {
            uint64_t cur = (rem << 32) | mag[i];
            mag[i] = (uint32_t)(cur / 1000000000u);
            rem = cur % 1000000000u;
        }
        while (!mag.empty() && mag.back() == 0) mag.pop_back();
        char buf[16];
        snprintf(buf, sizeof(buf), mag.empty() ? "%llu" : "%09llu", (unsigned long long)rem);
        digits.insert(0, buf);
    }
    return neg ? "-" + digits : digits;
}

struct ResultCopy {
    std::vector<uint32_t> pairs, offs;
    std::vector<uint64_t> exact;
    std::vector<int32_t> exp2;
    std::vector<double> f64;
    int limbs = 2;
};

void copy_result(const gi_result& r, ResultCopy& c) {
    c.pairs.assign(r.pairs, r.pairs + 2 * r.n_pairs);
    c.offs.assign(r.seg_offsets, r.seg_offsets + r.n_pairs + 1);
    c.limbs = r.exact_limbs;
    if (r.n_segments) {
        c.exact.assign(r.seg_exact, r.seg_exact + r.n_segments * 8 * (size_t)r.exact_limbs);
        c.f64.assign(r.seg_f64, r.seg_f64 + 6 * r.n_segments);
        if (r.seg_exp2) c.exp2.assign(r.seg_exp2, r.seg_exp2 + r.n_segments);
        else c.exp2.assign(r.n_segments, 0);
    } else {
        c.exact.clear();
        c.f64.clear();
        c.exp2.clear();
    }
}

void fill_stats(const gi_stats& s, uint32_t ntA, uint32_t ntB, ExecutionStats& st) {
    st.meshATotalNodes = ntA ? 2 * ntA - 1 : 0;
    st.meshBTotalNodes = ntB ? 2 * ntB - 1 : 0;
    st.totalPossiblePairs = (uint64_t)ntA * ntB;
    st.finalAabbCandidatePairs = s.candidates;
    st.confirmedGreenPairs = s.filter_green;
    st.confirmedYellowPairs = s.filter_yellow;
    st.traversalHits = s.traversal_hits;
    st.exactYellowYes = s.exact_yellow_yes;
    st.exactYellowNo = s.exact_yellow_no;
    st.hostResidualPairs = s.host_pairs;
    st.segments = s.segments;
    st.coplanarPairs = s.coplanar;
    st.degeneratePairs = s.degenerate;
    st.GPUTotalTime = s.ms_total;
    st.buildRefitMeshAMs = s.ms_build;
    st.buildRefitMeshBMs = 0.0;
    st.traverseMs = s.ms_traverse;
    st.filterMs = s.ms_filter;
    st.exactMs = s.ms_exact;
    st.outputMs = s.ms_output;
    st.hostMs = s.ms_host;
}

int2* malloc_pairs(const std::vector<uint32_t>& p, size_t n) {
    int2* out = (int2*)malloc(std::max<size_t>(n, 1) * sizeof(int2));
    if (!out) throw Error(GI_ERR_OUT_OF_MEMORY, "pair buffer allocation failed");
    for (size_t i = 0; i < n; ++i) out[i] = int2{(int)p[2 * i], (int)p[2 * i + 1]};
    return out;
}

}  // namespace

struct KernelBVHController::Impl {
    gi_context* ctx = nullptr;
    int device = 0;
    bool constructed = false;
    Point3 cA{0, 0, 0}, cB{0, 0, 0};
    uint32_t ntA = 0, ntB = 0;
    double3 rotA{0, 0, 0}, trA{0, 0, 0}, rotB{0, 0, 0}, trB{0, 0, 0};
    ResultCopy last;
};

KernelBVHController::KernelBVHController(int device) : m_impl(new Impl) {
    m_impl->device = device;
    gi_config cfg;
    memset(&cfg, 0, sizeof(cfg));
    cfg.device = device;
    check(gi_create(&cfg, &m_impl->ctx), nullptr);
}

KernelBVHController::~KernelBVHController() {
    if (m_impl && m_impl->ctx) gi_destroy(m_impl->ctx);
}

void KernelBVHController::construct(const Point3& centerA, const Point3& centerB, const double3* hVertsA,
                                    int numVertsA, const uint3* hIndicesA, int numTrianglesA, int /*levelA*/,
                                    const double3* hVertsB, int numVertsB, const uint3* hIndicesB, int numTrianglesB,
                                    int /*levelB*/, int /*leafThreshold*/, ExecutionStats& stats) {
    if (numTrianglesA <= 0 || numTrianglesB <= 0) return;  // reference: silent early return
    gi_context* c = m_impl->ctx;
    check(gi_set_mesh(c, GI_MESH_A, &hVertsA[0].x, (uint32_t)numVertsA, &hIndicesA[0].x, (uint32_t)numTrianglesA), c);
    check(gi_set_mesh(c, GI_MESH_B, &hVertsB[0].x, (uint32_t)numVertsB, &hIndicesB[0].x, (uint32_t)numTrianglesB), c);
    auto t0 = std::chrono::steady_clock::now();
    check(gi_build(c), c);
    auto t1 = std::chrono::steady_clock::now();
    m_impl->cA = centerA;
    m_impl->cB = centerB;
    m_impl->ntA = (uint32_t)numTrianglesA;
    m_impl->ntB = (uint32_t)numTrianglesB;
    m_impl->rotA = m_impl->trA = m_impl->rotB = m_impl->trB = double3{0, 0, 0};
    m_impl->constructed = true;
    stats.meshATotalNodes = 2 * m_impl->ntA - 1;
    stats.meshBTotalNodes = 2 * m_impl->ntB - 1;
    stats.buildRefitMeshAMs = std::chrono::duration<double, std::milli>(t1 - t0).count();
}

void KernelBVHController::construct(const double3* hVertsA, int numVertsA, const uint3* hIndicesA, int numTrianglesA,
                                    int levelA, const double3* hVertsB, int numVertsB, const uint3* hIndicesB,
                                    int numTrianglesB, int levelB, int leafThreshold, ExecutionStats& stats) {
    construct(centroid(hVertsA, numVertsA), centroid(hVertsB, numVertsB), hVertsA, numVertsA, hIndicesA,
              numTrianglesA, levelA, hVertsB, numVertsB, hIndicesB, numTrianglesB, levelB, leafThreshold, stats);
}

void KernelBVHController::runIntersectionPipeline(int /*batchMultiplier*/, int /*numberOfDualTreeSteps*/,
                                                  int /*activateAsyncDownload*/, int2*& outFinalExactPairs,
                                                  size_t& outFinalCount, ExecutionStats& stats, int /*gpuDouble*/) {
    if (!m_impl->constructed) return;
    gi_options o;
    memset(&o, 0, sizeof(o));
    o.want_segments = 1;
    gi_result r;
    check(gi_intersect(m_impl->ctx, &o, &r), m_impl->ctx);
    copy_result(r, m_impl->last);
    fill_stats(r.stats, m_impl->ntA, m_impl->ntB, stats);
    outFinalCount = r.n_pairs;
    gi_result_free(&r);
    outFinalExactPairs = malloc_pairs(m_impl->last.pairs, outFinalCount);
}

void KernelBVHController::cleanup() {
    gi_clear_device(m_impl->ctx);
    m_impl->constructed = false;
}

void KernelBVHController::clearGPU() { check(gi_clear_device(m_impl->ctx), m_impl->ctx); }

void KernelBVHController::reconstructGPU(ExecutionStats& stats, int, int) { reconstructGPU(stats); }

void KernelBVHController::reconstructGPU(ExecutionStats& stats) {
    auto t0 = std::chrono::steady_clock::now();
    check(gi_reconstruct_device(m_impl->ctx), m_impl->ctx);
    stats.buildRefitMeshAMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void KernelBVHController::setTransformBoth(double3 rotDegA, double3 transA, double3 rotDegB, double3 transB,
                                           float& timeGPU, float& timeCPU, float& timeTransformVerts,
                                           float& timeAssembleTris, float& timeGenBoxes) {
    auto t0 = std::chrono::steady_clock::now();
    gi_context* c = m_impl->ctx;
    double ra[3] = {rotDegA.x, rotDegA.y, rotDegA.z}, ta[3] = {transA.x, transA.y, transA.z};
    double rb[3] = {rotDegB.x, rotDegB.y, rotDegB.z}, tb[3] = {transB.x, transB.y, transB.z};
    double ca[3] = {m_impl->cA.x, m_impl->cA.y, m_impl->cA.z}, cb[3] = {m_impl->cB.x, m_impl->cB.y, m_impl->cB.z};
    check(gi_set_transform(c, GI_MESH_A, ra, ta, ca, snap_ ? 1 : 0), c);
    check(gi_set_transform(c, GI_MESH_B, rb, tb, cb, snap_ ? 1 : 0), c);
    check(gi_build(c), c);  // transform + boxes + refit
    m_impl->rotA = rotDegA; m_impl->trA = transA; m_impl->rotB = rotDegB; m_impl->trB = transB;
    float ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    timeGPU = ms;
    timeCPU = 0.f;
    timeTransformVerts = ms;
    timeAssembleTris = 0.f;
    timeGenBoxes = 0.f;
}

void KernelBVHController::setTranslation(double xB, double yB, double zB) {
    float a, b, c, d, e;
    setTransformBoth(m_impl->rotA, m_impl->trA, m_impl->rotB, double3{xB, yB, zB}, a, b, c, d, e);
}

Point3 KernelBVHController::getCenterA() const { return m_impl->cA; }
Point3 KernelBVHController::getCenterB() const { return m_impl->cB; }
bool KernelBVHController::isGPUAllocated() const { return m_impl->constructed; }

std::vector<IntersectionSegment> KernelBVHController::lastSegments() const {
    std::vector<IntersectionSegment> out;
    const ResultCopy& r = m_impl->last;
    for (size_t s = 0; s < r.f64.size() / 6; ++s) {
        const double* f = &r.f64[6 * s];
        out.push_back(IntersectionSegment{{f[0], f[1], f[2]}, {f[3], f[4], f[5]}});
    }
    return out;
}

std::vector<ExactSegment> KernelBVHController::lastExactSegments() const {
    std::vector<ExactSegment> out;
    const ResultCopy& r = m_impl->last;
    size_t npairs = r.offs.empty() ? 0 : r.offs.size() - 1;
    for (size_t i = 0; i < npairs; ++i)
        for (uint32_t s = r.offs[i]; s < r.offs[i + 1]; ++s) {
            const uint64_t* e = &r.exact[(size_t)s * 8 * r.limbs];
            ExactSegment x;
            x.a = (int)r.pairs[2 * i];
            x.b = (int)r.pairs[2 * i + 1];
            ExactRationalPoint* P[2] = {&x.p0, &x.p1};
            for (int k = 0; k < 2; ++k) {
                P[k]->x = to_decimal(e + (4 * k + 0) * r.limbs, r.limbs);
                P[k]->y = to_decimal(e + (4 * k + 1) * r.limbs, r.limbs);
                P[k]->z = to_decimal(e + (4 * k + 2) * r.limbs, r.limbs);
                P[k]->w = to_decimal(e + (4 * k + 3) * r.limbs, r.limbs);
                P[k]->exp2 = r.exp2.empty() ? 0 : r.exp2[s];
            }
            out.push_back(x);
        }
    return out;
}

void runGridForestIntersectionPipeline(const double3* hVertsA, int numVertsA, const uint3* hIndicesA,
                                       int numTrianglesA, int, const double3* hVertsB, int numVertsB,
                                       const uint3* hIndicesB, int numTrianglesB, int, int, int, int, int,
                                       ExecutionStats& stats, int2*& outFinalExactPairs, size_t& outFinalCount,
                                       Point3 centerA, Point3 centerB, double3 rotA, double3 transA, double3 rotB,
                                       double3 transB) {
    if (numTrianglesA <= 0 || numTrianglesB <= 0) return;  // reference: StandaloneBVHPipeline.cu:82-84
    KernelBVHController k;
    k.construct(centerA, centerB, hVertsA, numVertsA, hIndicesA, numTrianglesA, 0, hVertsB, numVertsB, hIndicesB,
                numTrianglesB, 0, 0, stats);
    float a, b, c, d, e;
    k.setTransformBoth(rotA, transA, rotB, transB, a, b, c, d, e);
    k.runIntersectionPipeline(0, 0, 0, outFinalExactPairs, outFinalCount, stats, 0);
}

std::vector<IntersectionSegment> computeExactIntersectionCurves(
    const double3* hVertsA, int numVertsA, const uint3* hIndicesA, int numTrianglesA, const double3* hVertsB,
    int numVertsB, const uint3* hIndicesB, int numTrianglesB, const std::vector<int2>& candidatePairs,
    const Point3& centerA, const double3& rotDegA, const double3& transA, const Point3& centerB,
    const double3& rotDegB, const double3& transB) {
    if (candidatePairs.empty()) return {};
    KernelBVHController k;
    ExecutionStats st;
    k.construct(centerA, centerB, hVertsA, numVertsA, hIndicesA, numTrianglesA, 0, hVertsB, numVertsB, hIndicesB,
                numTrianglesB, 0, 0, st);
    float a, b, c, d, e;
    k.setTransformBoth(rotDegA, transA, rotDegB, transB, a, b, c, d, e);
    int2* pairs = nullptr;
    size_t n = 0;
    k.runIntersectionPipeline(0, 0, 0, pairs, n, st, 0);
    free(pairs);
    // keep the segments of the requested pairs, in request order
    std::map<std::pair<int, int>, std::vector<IntersectionSegment>> by_pair;
    std::vector<ExactSegment> ex = k.lastExactSegments();
    std::vector<IntersectionSegment> f = k.lastSegments();
    for (size_t i = 0; i < ex.size(); ++i) by_pair[{ex[i].a, ex[i].b}].push_back(f[i]);
    std::vector<IntersectionSegment> out;
    for (const int2& p : candidatePairs) {
        auto it = by_pair.find({p.x, p.y});
        if (it != by_pair.end()) out.insert(out.end(), it->second.begin(), it->second.end());
    }
    return out;
}

}  // namespace gi_host
