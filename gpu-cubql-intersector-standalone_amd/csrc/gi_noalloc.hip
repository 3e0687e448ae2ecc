// gi_noalloc.hip -- gi_intersect_noalloc: the one-shot intersection of two
// meshes with no allocation of any kind by the library (no hipMalloc or
// hipFree, no pinned host memory, no streams or events, no host heap).
//
// Every intermediate (copies of the inputs, staged vertices, boxes, Morton
// keys, A's BVH, candidate and exact lists, result records, segment pool and
// the sorted output) is carved from one file-scope device array, `scratch`,
// of kScratchBytes. Each GPU the library runs on holds its own copy of the
// array from the code object's load until the process ends; its address on
// a device is looked up once (hipGetSymbolAddress, with that device current)
// and kept per device. Work runs on the null stream; the host reads the
// counters once (and the results) with synchronous copies into the caller's
// memory and the stack.
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind: counters are cleared by the first launch of each call. Calls that
// overlap on one GPU race (they use the same memory); callers serialise them.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "../../include/gi_intersector.h"
#include "gi_device.h"

namespace {

using namespace gi;

constexpr size_t kScratchBytes = (size_t)GI_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices];  // `scratch` on each device, resolved on first use there
std::mutex g_resolve;

// the device address of `scratch` on the current device (resolved once)
int scratch_on(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_resolve);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return GI_ERR_HIP;
        }
        g_scratch_ptr[device] = p;
    }
    *out = (unsigned char*)g_scratch_ptr[device];
    return GI_OK;
}

// Bump allocation from the scratch array (256-byte aligned pieces).
struct Carve {
    unsigned char* base;
    size_t off, cap;
    bool ok;
    template <class T>
    T* take(size_t n) {
        const size_t bytes = (std::max<size_t>(n, 1) * sizeof(T) + 255) & ~(size_t)255;
        if (!ok || off + bytes > cap) {
            ok = false;
            return nullptr;
        }
        T* p = (T*)(base + off);
        off += bytes;
        return p;
    }
};

int bits_for(uint32_t n) { return n <= 1 ? 1 : 32 - __builtin_clz(n - 1); }

// bounds of the vertices and whether they are all integers within 2^26 (the
// vertex stage's grid test on unmoved vertices)
bool bounds_and_grid(const double* xyz, uint32_t nv, double* lo, double* hi) {
    double l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool grid = true;
    for (size_t i = 0; i < nv; ++i)
        for (int k = 0; k < 3; ++k) {
            const double x = xyz[3 * i + k];
            l[k] = x < l[k] ? x : l[k];
            h[k] = x > h[k] ? x : h[k];
            grid = grid && x == rint(x) && fabs(x) <= 67108864.0;
        }
    for (int k = 0; k < 3; ++k) {
        lo[k] = l[k] <= h[k] ? l[k] : 0.0;
        hi[k] = l[k] <= h[k] ? h[k] : 0.0;
    }
    return grid;
}

struct MeshBufs {
    double* src;
    uint32_t* tri;
    double4* v64;
    int4* vi;
    float4* tbox;
    uint32_t *keys, *vals, *tk, *tv, *hist, *stmp;
    float4* nodes;  // A only
    uint2* range;
    float4* blocks;
    int axis;
    bool multi;  // multi-tile sort (first pass counted by the stage launch)
};

void carve_mesh(Carve& c, uint32_t nv, uint32_t nt, bool tree, MeshBufs& m) {
    m.src = c.take<double>(3 * (size_t)nv);
    m.tri = c.take<uint32_t>(3 * (size_t)nt);
    m.v64 = c.take<double4>(nv);
    m.vi = c.take<int4>(nv);
    m.tbox = c.take<float4>(2 * (size_t)nt);
    m.keys = c.take<uint32_t>(nt);
    m.vals = c.take<uint32_t>(nt);
    m.tk = c.take<uint32_t>(nt);
    m.tv = c.take<uint32_t>(nt);
    m.hist = c.take<uint32_t>(sort_hist_elems(nt));
    m.stmp = c.take<uint32_t>(scan_tmp_elems(sort_hist_elems(nt)));
    m.multi = radix_sort_takes_hist0(nt);
    m.nodes = tree ? c.take<float4>(2 * (size_t)(nt ? 2 * nt - 1 : 1)) : nullptr;
    m.range = tree ? c.take<uint2>(nt) : nullptr;
    m.blocks = tree ? c.take<float4>(block_levels_elems(nt)) : nullptr;
    m.axis = morton_axis_bits(nt);
}

// scratch bytes per unit of the arena scale X (X candidates, X/2 exact-list,
// residual and slow-list entries, X/4 result records and segments, and the
// sorted output of X/4 records), a bound used to size X from the space left
constexpr size_t kArenaBytesPerUnit = 8 + 8 + 4 + (8 + 8 + 8 * kSegPoolWords + 4 + 8 + 4 + 4 + 8 + 4 + 4 + 4 + 8 * kSegPoolWords +
                                               4 + 48 + 8) / 4 + 16;

#define NA_TRY(expr)                                              \
    do {                                                          \
        if ((expr) != hipSuccess) {                               \
            (void)hipGetLastError();                              \
            return GI_ERR_HIP;                                    \
        }                                                         \
    } while (0)

}  // namespace

extern "C" int gi_intersect_noalloc(int32_t device, const double* xyzA, uint32_t nvA, const uint32_t* triA,
                                    uint32_t ntA, const double* xyzB, uint32_t nvB, const uint32_t* triB, uint32_t ntB,
                                    int32_t want_segments, uint32_t* pairs, uint64_t pairs_cap, uint32_t* seg_offsets,
                                    uint64_t* seg_exact, int32_t* seg_exp2, uint64_t seg_cap, uint64_t* n_pairs,
                                    uint64_t* n_segments, int32_t* exact_limbs, gi_stats* stats) {
    if (!n_pairs || !n_segments || !exact_limbs || !seg_offsets || (pairs_cap && !pairs)) return GI_ERR_INVALID_ARGUMENT;
    if (want_segments && seg_cap && (!seg_exact || !seg_exp2)) return GI_ERR_INVALID_ARGUMENT;
    if ((nvA && !xyzA) || (ntA && !triA) || (nvB && !xyzB) || (ntB && !triB)) return GI_ERR_INVALID_ARGUMENT;
    if (device < 0 || device >= kMaxDevices) return GI_ERR_INVALID_ARGUMENT;
    *n_pairs = 0;
    *n_segments = 0;
    *exact_limbs = kSegLimbsGrid;
    seg_offsets[0] = 0;
    gi_stats st;
    memset(&st, 0, sizeof(st));
    if (stats) *stats = st;
    if (ntA > GI_NOALLOC_MAX_TRIANGLES || ntB > GI_NOALLOC_MAX_TRIANGLES || nvA > GI_NOALLOC_MAX_VERTICES ||
        nvB > GI_NOALLOC_MAX_VERTICES)
        return GI_ERR_TOO_LARGE;
    for (size_t i = 0; i < 3 * (size_t)ntA; ++i)
        if (triA[i] >= nvA) return GI_ERR_INVALID_ARGUMENT;
    for (size_t i = 0; i < 3 * (size_t)ntB; ++i)
        if (triB[i] >= nvB) return GI_ERR_INVALID_ARGUMENT;
    if (ntA == 0 || ntB == 0) return GI_OK;  // reference: an empty mesh gives no pairs
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device >= ndev) {
        (void)hipGetLastError();
        return GI_ERR_NO_DEVICE;
    }
    NA_TRY(hipSetDevice(device));
    unsigned char* base = nullptr;
    if (int rc = scratch_on(device, &base)) return rc;

    // source bounds (the Morton frames) and the grid classification
    Xform xf[2];
    memset(xf, 0, sizeof(xf));
    int src_grid[2];
    const double* xyz[2] = {xyzA, xyzB};
    const uint32_t nv[2] = {nvA, nvB}, nt[2] = {ntA, ntB};
    for (int s = 0; s < 2; ++s) {
        xf[s].m[0][0] = xf[s].m[1][1] = xf[s].m[2][2] = 1.0;
        xf[s].identity = 1;
        src_grid[s] = bounds_and_grid(xyz[s], nv[s], xf[s].lo, xf[s].hi) ? 1 : 0;
    }

    // the scratch layout: both meshes, then the arenas in what is left
    Carve c{base, 0, kScratchBytes, true};
    uint32_t* counters = c.take<uint32_t>(CNT_COUNT + 2);
    Xform* d_xf = c.take<Xform>(2);
    MeshBufs M[2];
    carve_mesh(c, nvA, ntA, true, M[0]);
    carve_mesh(c, nvB, ntB, false, M[1]);
    if (!c.ok) return GI_ERR_TOO_LARGE;
    const size_t left = c.cap - c.off;
    uint64_t X = (left / kArenaBytesPerUnit) & ~(uint64_t)1023;
    X = std::min<uint64_t>(X, 0xFFFFFFF0ull & ~(uint64_t)1023);
    if (X < 4096) return GI_ERR_TOO_LARGE;
    const uint32_t cap_cand = (uint32_t)X, cap_exact = (uint32_t)(X / 2), cap_rec = (uint32_t)(X / 4),
                   cap_seg = (uint32_t)(X / 4);
    uint2* cand = c.take<uint2>(cap_cand);
    uint2* exact = c.take<uint2>(cap_exact);
    uint2* resid = c.take<uint2>(cap_exact);
    uint2* slow = c.take<uint2>(cap_exact);
    uint64_t* rec_key = c.take<uint64_t>(cap_rec);
    uint2* rec_seg = c.take<uint2>(cap_rec);
    uint64_t* seg_pool = c.take<uint64_t>(kSegPoolWords * (size_t)cap_seg);
    int32_t* seg_exp = c.take<int32_t>(cap_seg);
    uint2* out_pairs = c.take<uint2>(cap_rec);
    uint32_t* order = c.take<uint32_t>(cap_rec);
    uint32_t* offs = c.take<uint32_t>((size_t)cap_rec + 1);
    uint64_t* tkeys = c.take<uint64_t>(cap_rec);
    uint32_t* tvals = c.take<uint32_t>(cap_rec);
    uint32_t* nseg = c.take<uint32_t>(cap_rec);
    uint32_t* rhist = c.take<uint32_t>(sort_hist_elems(cap_rec));
    uint32_t* rstmp = c.take<uint32_t>(scan_tmp_elems(std::max(sort_hist_elems(cap_rec), (size_t)cap_rec + 1)));
    uint64_t* segs = c.take<uint64_t>(kSegPoolWords * (size_t)cap_seg);
    int32_t* sexp = c.take<int32_t>(cap_seg);
    double* f64 = c.take<double>(6 * (size_t)cap_seg);
    if (!c.ok) return GI_ERR_TOO_LARGE;

    hipStream_t s = 0;  // the null stream
    for (int k = 0; k < 2; ++k) {
        if (nv[k]) NA_TRY(hipMemcpy(M[k].src, xyz[k], 24 * (size_t)nv[k], hipMemcpyHostToDevice));
        NA_TRY(hipMemcpy(M[k].tri, k ? triB : triA, 12 * (size_t)nt[k], hipMemcpyHostToDevice));
    }
    NA_TRY(hipMemcpy(d_xf, xf, sizeof(xf), hipMemcpyHostToDevice));
    uint32_t* grid = counters + CNT_COUNT;
    launch_query_init(counters, CNT_COUNT, grid, 3u, nullptr, 0u, s);
    SortJobU32 jobs[2];
    StageJob sj[2];
    for (int k = 0; k < 2; ++k) {
        MeshBufs& m = M[k];
        sj[k] = StageJob{m.src, nv[k], m.tri, nt[k], d_xf + k, m.v64, m.vi, grid + k, m.axis, m.tbox, m.keys, m.vals,
                         m.multi ? m.hist : nullptr};
        jobs[k] = SortJobU32{m.keys, m.vals, m.tk, m.tv, nt[k], 3 * m.axis, m.multi, m.hist, m.stmp};
    }
    launch_stage_meshes(sj, 2, s);
    int in_tmp[2] = {0, 0};
    radix_sort_u32_pair(jobs, 2, s, in_tmp);
    BuildJob<uint32_t> bj;
    bj.keys = in_tmp[0] ? M[0].tk : M[0].keys;
    bj.prims = in_tmp[0] ? M[0].tv : M[0].vals;
    bj.n = ntA;
    bj.boxes = M[0].tbox;
    bj.nodes = M[0].nodes;
    bj.range = M[0].range;
    bj.blocks = M[0].blocks;
    launch_build_bvh(&bj, 1, kBuildChunk, s, nullptr, nullptr, 0);
    MeshDev A, B;
    MeshDev* D[2] = {&A, &B};
    for (int k = 0; k < 2; ++k) {
        D[k]->nv = nv[k];
        D[k]->nt = nt[k];
        D[k]->src = M[k].src;
        D[k]->v64 = M[k].v64;
        D[k]->vi = M[k].vi;
        D[k]->tri = M[k].tri;
        D[k]->nodes = M[k].nodes;
        D[k]->order = in_tmp[k] ? M[k].tv : M[k].vals;
        D[k]->tbox = M[k].tbox;
    }
    const int bitsA = bits_for(ntA), bitsB = bits_for(ntB);
    const uint32_t out_slots = std::min(output_small_max(), cap_seg);
    int kinds = (src_grid[0] && src_grid[1]) ? kExactGrid : kExactWide;  // unmoved vertices: known on the host
    uint32_t hc[CNT_COUNT + 2];
    for (int pass = 0;; ++pass) {
        if (pass) launch_query_init(counters, CNT_COUNT, grid, 0u, nullptr, 0u, s);  // queries only; the trees are built
        launch_traverse(A, B, nullptr, 0, nullptr, cand, cap_cand, counters, 2, s);
        launch_filter(A, B, cand, cap_cand, grid, exact, cap_exact, counters, 2048, s);
        launch_exact(A, B, exact, cap_exact, grid, bitsB, rec_key, rec_seg, cap_rec, seg_pool, seg_exp, cap_seg, cand,
                     cap_cand, slow, cap_exact, resid, cap_exact, counters, 1024, kinds, s);
        launch_output_small(rec_key, rec_seg, counters, cap_rec, cap_seg, bitsA + bitsB, bitsB, want_segments != 0,
                            out_pairs, order, offs, segs, sexp, f64, out_slots, seg_pool, seg_exp, grid, nullptr, s);
        NA_TRY(hipGetLastError());
        NA_TRY(hipMemcpy(hc, counters, sizeof(hc), hipMemcpyDeviceToHost));
        const bool is_grid = hc[CNT_COUNT] && hc[CNT_COUNT + 1];
        if (kinds & (is_grid ? kExactGrid : kExactWide)) break;
        if (pass) return GI_ERR_INTERNAL;
        kinds = kExactGrid | kExactWide;  // the host classification did not hold: both stages
    }
    const bool is_grid = hc[CNT_COUNT] && hc[CNT_COUNT + 1];
    st.traversal_hits = hc[CNT_TRAVERSAL_HITS];
    st.candidates = hc[CNT_CANDIDATES];
    st.filter_no = hc[CNT_FILTER_NO];
    st.filter_green = hc[CNT_FILTER_GREEN];
    st.filter_yellow = hc[CNT_FILTER_YELLOW];
    st.filter_predicates = hc[CNT_FILTER_PRED];
    st.exact_predicates = hc[CNT_EXACT_PRED];
    st.exact_pairs = hc[CNT_EXACT_LIST] - hc[CNT_RESIDUAL];
    st.exact_yellow_yes = hc[CNT_EXACT_YELLOW_YES];
    st.exact_yellow_no = hc[CNT_EXACT_YELLOW_NO];
    st.coplanar = hc[CNT_COPLANAR];
    st.degenerate = hc[CNT_DEGENERATE];
    st.host_pairs = hc[CNT_RESIDUAL];
    st.grid = is_grid ? 1 : 0;
    const uint32_t R = hc[CNT_RESULTS], S = want_segments ? hc[CNT_SEGMENTS] : 0;
    st.pairs = R;
    st.segments = S;
    if (stats) *stats = st;
    if (hc[CNT_OVERFLOW] || hc[CNT_TRAVERSAL_HITS] > cap_cand || hc[CNT_EXACT_LIST] > cap_exact ||
        hc[CNT_RESULTS] > cap_rec || hc[CNT_SEGMENTS] > cap_seg)
        return GI_ERR_TOO_LARGE;  // beyond what the scratch array holds
    if (hc[CNT_UNSOUND]) return GI_ERR_INTERNAL;
    if (hc[CNT_RESIDUAL]) return GI_ERR_TOO_LARGE;  // pairs that need the host residual stage
    *n_pairs = R;
    *n_segments = S;
    const int L = is_grid ? kSegLimbsGrid : kSegLimbsWide;
    *exact_limbs = L;
    if (R > pairs_cap || S > seg_cap) return GI_ERR_TOO_LARGE;  // the caller's buffers: counts reported
    const bool done = output_small_fits(R, bitsA + bitsB);
    if (!done) {  // the sorted output of a larger set, after the read-back
        uint64_t* keys = rec_key;
        uint32_t* ord = order;
        launch_iota(ord, R, s);
        if (radix_sort_u64(keys, ord, tkeys, tvals, R, bitsA + bitsB, rhist, rstmp, s)) {
            keys = tkeys;
            ord = tvals;
        }
        launch_gather_pairs(keys, ord, rec_seg, R, bitsB, out_pairs, nseg, s);
        if (!want_segments) launch_fill_u32(nseg, 0u, R, s);
        exclusive_scan_u32(nseg, offs, R, rstmp, offs + R, s);
        if (want_segments)
            launch_gather_segments(ord, rec_seg, offs, R, seg_pool, seg_exp, L, segs, sexp, f64, s);
    } else if (S > out_slots) {
        launch_gather_segments(order, rec_seg, offs, R, seg_pool, seg_exp, L, segs, sexp, f64, s);
    }
    NA_TRY(hipGetLastError());
    if (R) NA_TRY(hipMemcpy(pairs, out_pairs, 8 * (size_t)R, hipMemcpyDeviceToHost));
    NA_TRY(hipMemcpy(seg_offsets, offs, 4 * ((size_t)R + 1), hipMemcpyDeviceToHost));
    if (S) {
        NA_TRY(hipMemcpy(seg_exact, segs, 64 * (size_t)L * S, hipMemcpyDeviceToHost));
        if (is_grid) memset(seg_exp2, 0, 4 * (size_t)S);
        else NA_TRY(hipMemcpy(seg_exp2, sexp, 4 * (size_t)S, hipMemcpyDeviceToHost));
    }
    return GI_OK;
}
