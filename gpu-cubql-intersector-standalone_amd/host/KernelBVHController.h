// KernelBVHController.h -- C++ drop-in for the reference's GPUIntersector
// surface, implemented over the C ABI (include/gi_intersector.h).
//
// Mirrors, method for method and argument for argument:
//   class KernelBVHController      src/GPUIntersector/KernelBVHController.h:36-95
//   runGridForestIntersectionPipeline
//                                  src/GPUIntersector/StandaloneBVHPipeline.h:27-42
//   computeExactIntersectionCurves src/CPU/MeshIntersectionCurves.h:24-29
//
// Differences a caller sees:
//   - CGAL's Mesh and Point3 are replaced by plain arrays and a POD Point3;
//     the reference passes both the CGAL mesh and the same data as arrays.
//   - BVH tuning arguments (levels, leafThreshold, batchMultiplier,
//     numberOfDualTreeSteps) are accepted and recorded but the MI355X build
//     has no level cut or dual-tree step count to tune; gpuDouble (predicate
//     mode) is accepted and ignored: every mode ends in the exact decision.
//   - Failures throw gi_host::Error instead of exit(EXIT_FAILURE)
//     (reference: src/single_tree/SingleTreeBatchCommon.h:13-23); an empty
//     mesh still returns early with the outputs untouched
//     (KernelBVHController.cu:461-462).
//   - The pair buffer is malloc'ed and owned by the caller, who std::free's
//     it, as in the reference (apps/mainApplication.cpp:793-796).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace gi_host {

struct double3 { double x, y, z; };
struct uint3 { uint32_t x, y, z; };
struct int2 { int x, y; };
struct Point3 { double x, y, z; };

// reference: src/CPU/MeshIntersectionCurves.h:9-17 (endpoints as doubles);
// the exact rational endpoints are available through exactSegments()
struct ExactPoint3POD { double x, y, z; };
struct IntersectionSegment { ExactPoint3POD p0, p1; };

// Exact homogeneous endpoint (x/w, y/w, z/w) * 2^exp2 as decimal strings.
struct ExactRationalPoint {
    std::string x, y, z, w;
    int exp2;
};
struct ExactSegment {
    int a, b;  // face pair
    ExactRationalPoint p0, p1;
};

// reference: src/common/ExecutionStats.h:12-45 (fields this build fills)
struct ExecutionStats {
    uint32_t meshATotalNodes = 0;
    uint32_t meshBTotalNodes = 0;
    uint64_t totalPossiblePairs = 0;
    double buildRefitMeshAMs = 0.0;  // both builds (A and B are built in one stream)
    double buildRefitMeshBMs = 0.0;
    double GPUTotalTime = 0.0;
    uint64_t finalAabbCandidatePairs = 0;
    size_t confirmedGreenPairs = 0;   // certified by the FP64 filter
    size_t confirmedYellowPairs = 0;  // deferred to the exact stage
    // stage counters of this build's cascade
    uint64_t traversalHits = 0;
    uint64_t exactYellowYes = 0, exactYellowNo = 0;
    uint64_t hostResidualPairs = 0;
    uint64_t segments = 0, coplanarPairs = 0, degeneratePairs = 0;
    double traverseMs = 0.0, filterMs = 0.0, exactMs = 0.0, outputMs = 0.0, hostMs = 0.0;
};

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& msg) : std::runtime_error(msg), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

class KernelBVHController {
public:
    explicit KernelBVHController(int device = 0);
    ~KernelBVHController();
    KernelBVHController(const KernelBVHController&) = delete;
    KernelBVHController& operator=(const KernelBVHController&) = delete;

    void construct(const Point3& centerA, const Point3& centerB, const double3* hVertsA, int numVertsA,
                   const uint3* hIndicesA, int numTrianglesA, int levelA, const double3* hVertsB, int numVertsB,
                   const uint3* hIndicesB, int numTrianglesB, int levelB, int leafThreshold, ExecutionStats& stats);
    // centres = vertex centroids
    void construct(const double3* hVertsA, int numVertsA, const uint3* hIndicesA, int numTrianglesA, int levelA,
                   const double3* hVertsB, int numVertsB, const uint3* hIndicesB, int numTrianglesB, int levelB,
                   int leafThreshold, ExecutionStats& stats);

    void runIntersectionPipeline(int batchMultiplier, int numberOfDualTreeSteps, int activateAsyncDownload,
                                 int2*& outFinalExactPairs, size_t& outFinalCount, ExecutionStats& stats,
                                 int gpuDouble);

    void cleanup();
    void clearGPU();
    void reconstructGPU(ExecutionStats& stats, int a, int b);
    void reconstructGPU(ExecutionStats& stats);

    // Euler angles in degrees (R = Rz Ry Rx) about each mesh's centre, then
    // the translation; absolute (always applied to the constructed vertices).
    void setTransformBoth(double3 rotDegA, double3 transA, double3 rotDegB, double3 transB, float& timeGPU,
                          float& timeCPU, float& timeTransformVerts, float& timeAssembleTris, float& timeGenBoxes);
    void setTranslation(double xB, double yB, double zB);

    Point3 getCenterA() const;
    Point3 getCenterB() const;
    bool isGPUAllocated() const;

    // Segments of the last runIntersectionPipeline, in pair order.
    std::vector<IntersectionSegment> lastSegments() const;
    std::vector<ExactSegment> lastExactSegments() const;

    // Snap moved vertices to integers (keeps grid meshes on the GPU exact stage).
    void setSnapToGrid(bool snap) { snap_ = snap; }

private:
    struct Impl;
    std::unique_ptr<Impl> m_impl;
    bool snap_ = false;
};

// reference: extern "C" runGridForestIntersectionPipeline (per-call build).
// gridLevel*, batchMultiplier, steps, mode and leafThreshold are accepted
// and ignored; see the class comment.
void runGridForestIntersectionPipeline(const double3* hVertsA, int numVertsA, const uint3* hIndicesA,
                                       int numTrianglesA, int maxCellSizeA, const double3* hVertsB, int numVertsB,
                                       const uint3* hIndicesB, int numTrianglesB, int maxCellSizeB,
                                       int batchMultiplier, int numberOfDualTreeSteps, int gpuPredicateMode,
                                       int leafThreshold, ExecutionStats& stats, int2*& outFinalExactPairs,
                                       size_t& outFinalCount, Point3 centerA, Point3 centerB, double3 rotA,
                                       double3 transA, double3 rotB, double3 transB);

// reference: computeExactIntersectionCurves (src/CPU/MeshIntersectionCurves.h:24-29)
std::vector<IntersectionSegment> computeExactIntersectionCurves(
    const double3* hVertsA, int numVertsA, const uint3* hIndicesA, int numTrianglesA, const double3* hVertsB,
    int numVertsB, const uint3* hIndicesB, int numTrianglesB, const std::vector<int2>& candidatePairs,
    const Point3& centerA, const double3& rotDegA, const double3& transA, const Point3& centerB,
    const double3& rotDegB, const double3& transB);

}  // namespace gi_host
