// Drives the C++ facade (KernelBVHController) the way the reference's
// application drives its controller (apps/mainApplication.cpp:723-796):
// construct -> runIntersectionPipeline -> setTransformBoth -> run again, then
// the per-call free function runGridForestIntersectionPipeline
// (src/GPUIntersector/StandaloneBVHPipeline.h:27-42) with the same motion of B.
// Meshes come from raw files: uint32 nv, uint32 nt, nv*3 doubles, nt*3 uint32.
// Arguments: meshA meshB out [rx ry rz tx ty tz [snap]] (snap: 1 rounds the
// moved vertices to integers, 0 keeps the moved doubles; default 1).
// Output: one block per run: "run <n>", then "pair a b" lines, then
// "seg a b x0 y0 z0 w0 x1 y1 z1 w1 exp2" lines (exact decimal integers); the
// free function's run has pairs only.
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "KernelBVHController.h"

using namespace gi_host;

static bool read_mesh(const char* path, std::vector<double3>& v, std::vector<uint3>& t) {
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    uint32_t nv, nt;
    if (fread(&nv, 4, 1, f) != 1 || fread(&nt, 4, 1, f) != 1) { fclose(f); return false; }
    v.resize(nv);
    t.resize(nt);
    bool ok = fread(v.data(), sizeof(double3), nv, f) == nv && fread(t.data(), sizeof(uint3), nt, f) == nt;
    fclose(f);
    return ok;
}

static void dump(FILE* o, KernelBVHController& k, int2* pairs, size_t n, const ExecutionStats& st) {
    fprintf(o, "run %zu candidates %llu\n", n, (unsigned long long)st.finalAabbCandidatePairs);
    for (size_t i = 0; i < n; ++i) fprintf(o, "pair %d %d\n", pairs[i].x, pairs[i].y);
    for (const ExactSegment& s : k.lastExactSegments())
        fprintf(o, "seg %d %d %s %s %s %s %s %s %s %s %d\n", s.a, s.b, s.p0.x.c_str(), s.p0.y.c_str(), s.p0.z.c_str(),
                s.p0.w.c_str(), s.p1.x.c_str(), s.p1.y.c_str(), s.p1.z.c_str(), s.p1.w.c_str(), s.p0.exp2);
}

int main(int argc, char** argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s meshA meshB out [rx ry rz tx ty tz]\n", argv[0]);
        return 2;
    }
    std::vector<double3> vA, vB;
    std::vector<uint3> tA, tB;
    if (!read_mesh(argv[1], vA, tA) || !read_mesh(argv[2], vB, tB)) {
        fprintf(stderr, "cannot read meshes\n");
        return 2;
    }
    double3 rot{0, 0, 0}, tr{0, 0, 0};
    if (argc >= 10) {
        rot = double3{atof(argv[4]), atof(argv[5]), atof(argv[6])};
        tr = double3{atof(argv[7]), atof(argv[8]), atof(argv[9])};
    }
    const bool snap = argc >= 11 ? atoi(argv[10]) != 0 : true;
    FILE* o = fopen(argv[3], "w");
    if (!o) return 2;
    try {
        KernelBVHController k(0);
        k.setSnapToGrid(snap);
        ExecutionStats st;
        k.construct(vA.data(), (int)vA.size(), tA.data(), (int)tA.size(), 12, vB.data(), (int)vB.size(), tB.data(),
                    (int)tB.size(), 12, 20, st);
        int2* pairs = nullptr;
        size_t n = 0;
        k.runIntersectionPipeline(1, 12, 0, pairs, n, st, 1);
        dump(o, k, pairs, n, st);
        free(pairs);
        float a, b, c, d, e;
        k.setTransformBoth(double3{0, 0, 0}, double3{0, 0, 0}, rot, tr, a, b, c, d, e);
        k.runIntersectionPipeline(1, 12, 0, pairs, n, st, 1);
        dump(o, k, pairs, n, st);
        free(pairs);
        Point3 cb = k.getCenterB();
        fprintf(o, "centerB %.17g %.17g %.17g\n", cb.x, cb.y, cb.z);
        if (!snap) {  // the free function keeps the moved doubles, as the reference does
            ExecutionStats fs;
            int2* fp = nullptr;
            size_t fn = 0;
            runGridForestIntersectionPipeline(vA.data(), (int)vA.size(), tA.data(), (int)tA.size(), 12, vB.data(),
                                              (int)vB.size(), tB.data(), (int)tB.size(), 12, 1, 12, 1, 20, fs, fp,
                                              fn, k.getCenterA(), cb, double3{0, 0, 0}, double3{0, 0, 0}, rot, tr);
            fprintf(o, "run %zu candidates %llu\n", fn, (unsigned long long)fs.finalAabbCandidatePairs);
            for (size_t i = 0; i < fn; ++i) fprintf(o, "pair %d %d\n", fp[i].x, fp[i].y);
            free(fp);
        }
    } catch (const Error& e) {
        fprintf(stderr, "facade error %d: %s\n", e.status(), e.what());
        fclose(o);
        return 1;
    }
    fclose(o);
    return 0;
}
