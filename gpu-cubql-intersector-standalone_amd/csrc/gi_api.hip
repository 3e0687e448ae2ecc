// gi_api.hip -- gi_context and the C ABI declared in include/gi_intersector.h.
//
// One context owns one HIP stream, both meshes (host copy + device state),
// the candidate/exact/result arenas and the events used for stage timings.
// A query runs the whole device pipeline asynchronously and synchronises
// once, to read the counters (capacity check, grid/non-grid branch).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gi_intersector.h"
#include "gi_device.h"
#include "host_exact.h"

namespace {

using namespace gi;

struct GiError {
    int status;
    std::string msg;
};

#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw GiError{e_ == hipErrorOutOfMemory ? GI_ERR_OUT_OF_MEMORY : GI_ERR_HIP,               \
                          std::string(#expr) + ": " + hipGetErrorString(e_)};                          \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* ensure(size_t n) {
        size_t need = std::max<size_t>(n, 1) * sizeof(T);
        if (bytes < need) {
            release();
            HIP_TRY(hipMalloc(&p, need));
            bytes = need;
        }
        return (T*)p;
    }
    template <class T>
    T* get() const { return (T*)p; }
};

struct MeshSlot {
    // host copies of the source arrays, made only when needed (clear /
    // reconstruct, the host residual stage): the upload reads the caller's
    // arrays directly
    std::vector<double> h_xyz;
    std::vector<uint32_t> h_tri;
    bool host_valid = false;
    uint32_t nv = 0, nt = 0;
    bool set = false;
    bool on_device = false;  // device buffers hold the mesh
    bool built = false;      // BVH topology is valid
    bool dirty = true;       // vertices/boxes/BVH boxes need an update
    // Mesh B's BVH is walked by nothing (its triangles, in Morton order with
    // their boxes, are the queries against A's tree): B builds no tree unless
    // it is shared into an A slot (gi_share_mesh), which needs the whole tree.
    bool full_tree = false;
    Xform xf;
    int sorted_in_tmp = 0;   // the sorted Morton keys live in (tkeys, tvals)
    bool borrowed = false;     // device buffers belong to another context (gi_share_mesh)
    // batch ("forest"): `forest` meshes stored back to back, indices global;
    // mesh m owns vertices [h_voff[m], h_voff[m+1]) and triangles
    // [h_toff[m], h_toff[m+1]); tiles = (mesh, first element) pairs
    uint32_t forest = 0;
    std::vector<uint32_t> h_voff, h_toff, h_vtiles, h_ttiles, h_vtile0;
    int morton_axis = 10;      // Morton bits per axis of the kept sorted keys
    double src_lo[3] = {0.0, 0.0, 0.0}, src_hi[3] = {0.0, 0.0, 0.0};  // bounds of the source vertices
    int src_grid = -1;  // the source vertices are all integers within 2^26 (1), not (0), unknown (-1)
    // the last staging wrote the moved copies (v64, vi); without a move (and
    // without a snap that changes them) the kernels read the source vertices
    bool moved_copy = true;
    uint64_t generation = 0;   // bumped whenever the slot's mesh or buffers are replaced
    DevBuf src, tri, v64, vi, tbox, keys, vals, tkeys, tvals, nodes, range, blocks;
    DevBuf voff, toff, vtiles, ttiles, vtile0, fbounds, roots;
    // the source vertices as integers (an owned slot whose source is on the
    // grid): made on the first staging after an upload, dropped with the
    // source; the unmoved mesh's kernels read it instead of the doubles
    DevBuf vsrc;
    bool vsrc_ok = false;
    void release() {
        ++generation;
        DevBuf* forest_bufs[] = {&voff, &toff, &vtiles, &ttiles, &vtile0, &fbounds, &roots};
        for (DevBuf* b : forest_bufs) b->release();
        vsrc.release();
        vsrc_ok = false;
        DevBuf* all[] = {&src, &tri, &v64, &vi, &tbox, &keys, &vals, &tkeys, &tvals, &nodes, &range, &blocks};
        for (DevBuf* b : all) {
            if (borrowed) { b->p = nullptr; b->bytes = 0; }
            else b->release();
        }
        if (borrowed) { borrowed = false; set = false; }
        moved_copy = true;
        on_device = false;
        full_tree = false;
        built = false;
        dirty = true;
    }
};

Xform identity_xform() {
    Xform x;
    memset(&x, 0, sizeof(x));
    x.m[0][0] = x.m[1][1] = x.m[2][2] = 1.0;
    x.identity = 1;
    return x;
}

constexpr double kGridBound = 67108864.0;  // 2^26: the grid stage's coordinate bound

int bits_for(uint32_t n) { return n <= 1 ? 1 : 32 - __builtin_clz(n - 1); }

// What the vertex stage's grid test will find for a slot, when the host can
// tell: 1 on the grid, 0 off it, -1 unknown. Unmoved vertices keep their
// source classification; snapped ones are integers, within the bound when
// the moved corners of the source box are (margin for the rounding of the
// map). An unsnapped move whose matrix is not a signed permutation is taken
// to leave the grid (a rotation by any other angle maps integers to
// non-integers but in rare rational cases; a wrong guess only costs the
// query a second run with both exact stages); other moves are left to the
// device.
int predicted_grid(const MeshSlot& m) {
    if (m.src_grid < 0) return -1;
    if (!m.xf.identity && !m.xf.snap) {
        bool perm = true;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                const double v = m.xf.m[r][c];
                perm = perm && (v == 0.0 || v == 1.0 || v == -1.0);
            }
        return perm ? -1 : 0;
    }
    if (m.xf.identity) {
        if (!m.xf.snap || m.src_grid == 1) return m.src_grid;
        const double r = std::max(std::max(fabs(m.src_lo[0]), fabs(m.src_hi[0])),
                                  std::max(std::max(fabs(m.src_lo[1]), fabs(m.src_hi[1])),
                                           std::max(fabs(m.src_lo[2]), fabs(m.src_hi[2]))));
        return r <= kGridBound - 2.0 ? 1 : -1;
    }
    if (!m.xf.snap) return -1;
    for (int c = 0; c < 8; ++c) {
        const double p[3] = {(c & 1) ? m.src_hi[0] : m.src_lo[0], (c & 2) ? m.src_hi[1] : m.src_lo[1],
                             (c & 4) ? m.src_hi[2] : m.src_lo[2]};
        for (int r = 0; r < 3; ++r) {
            const double v = m.xf.m[r][0] * p[0] + m.xf.m[r][1] * p[1] + m.xf.m[r][2] * p[2] + m.xf.m[r][3];
            if (!(fabs(v) <= kGridBound - 2.0)) return -1;
        }
    }
    return 1;
}

}  // namespace

struct gi_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // mesh B's build runs on `side`, concurrently with mesh A's on `stream`
    // (fork/join by events); build_stream[slot] is the stream of that slot
    hipStream_t side = nullptr;
    hipStream_t build_stream[2] = {nullptr, nullptr};
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    DevBuf bhist[2], bscan[2], bpart[2];  // per-slot build scratch
    SortJobU32 pending_sort[2] = {};       // Morton sorts queued by the slot stages (n = 0: none)
    MsdSortJob pending_msd[2] = {};        // the same as bucket sorts (keys null: none)
    DevBuf msd;                            // bucket-sort words of both slots (2 kMsdWords)
    std::string err;
    int host_threads = 0;
    uint32_t cand_reserve = 0;
    MeshSlot mesh[2];
    DevBuf counters, hist, scan_tmp, partials;
    DevBuf inspect;  // partial bounds and the MeshInspect of the last upload
    DevBuf cand, exact, rec_key, rec_seg, seg_pool, seg_exp, resid;
    DevBuf slow;  // pairs the one-lane exact passes leave to the general ones (as many as `exact`)
    uint32_t cap_cand = 0, cap_exact = 0, cap_rec = 0, cap_seg = 0;
    DevBuf order, tkeys, tvals, pairs, nseg, offs, segs, f64, exp2, pair_off;
    std::vector<uint32_t> h_pair_off;  // per mesh pair result offsets of the last batch query
    gi_tuning tune = {2, 2048, 1024, (int32_t)kBuildChunk};  // launch shapes (gi_autotune)
    hipEvent_t ev[10] = {};  // [8..9]: the grid exact stage's passes after the read-back
    // pinned host staging: per-slot transforms (copied to xf_dev before the
    // vertex kernels) and the counter read-back; graph replays read/write them
    Xform* h_xf = nullptr;
    uint32_t* h_counters = nullptr;
    uint32_t* d_counters_host = nullptr;  // h_counters as kernels address it
    DevBuf xf_dev;
    // what an eager stage last copied into xf_dev[slot]: an unchanged
    // transform is not sent again (captures and replays clear the flag)
    Xform xf_last[2] = {};
    bool xf_current[2] = {false, false};
    // captured launch sequence of a query up to the counter read-back, valid
    // while graph_sig (sizes, flags, device pointers) is unchanged
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint64_t> graph_sig, graph_sig_seen;
    bool graph_off = false;  // capture failed once: queries stay eager
    int exact_kinds = kExactGrid | kExactWide;  // exact stages the last presync launched
    bool exact_both = false;  // this query launches both exact stages (a prediction failed)
    bool capturing = false;  // a capture is open: timing events become graph nodes
    bool graph_chunk_timed = false;  // what the capture's build left in chunk_timed/leaves
    uint32_t graph_chunk_leaves = 0;
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_sig.clear();
    }
    // pinned staging of large batch uploads (allocated on the first one,
    // kept until destroy): per worker two buffers, a stream and two events
    struct Stager {
        std::vector<void*> buf;
        std::vector<hipStream_t> st;
        std::vector<hipEvent_t> ev;
        int workers = 0;
    } stager;
    hipEvent_t ev_chunk[4] = {};  // [0..1] bracket the last k_build_chunks launch
    bool chunk_timed = false;     // that launch ran in the current call
    hipEvent_t ev_stage[2] = {};  // bracket the last staging launch of single meshes
    bool stage_timed = false;     // that launch ran, timed, in the current call
    bool graph_stage_timed = false;
    uint32_t chunk_leaves = 0;    // leaves it covered (both trees when both were built)
    void release_device() {
        drop_graph();
        graph_sig_seen.clear();
        xf_dev.release();
        xf_current[0] = xf_current[1] = false;
        for (auto& m : mesh) m.release();
        DevBuf* all[] = {&counters, &hist, &scan_tmp, &partials, &inspect, &cand, &exact, &rec_key, &rec_seg, &seg_pool,
                         &seg_exp, &resid, &order, &tkeys, &tvals, &pairs, &nseg, &offs, &segs, &f64, &exp2, &pair_off,
                         &bhist[0], &bhist[1], &bscan[0], &bscan[1], &bpart[0], &bpart[1], &msd};
        for (DevBuf* b : all) b->release();
        cap_cand = cap_exact = cap_rec = cap_seg = 0;
    }
};

namespace {

struct HostResultStore {
    std::vector<uint32_t> pairs, offs;
    std::vector<uint64_t> seg;
    std::vector<int32_t> exp2;
    std::vector<double> f64;
};

int fail(gi_context* ctx, const GiError& e) {
    if (ctx) ctx->err = e.msg;
    return e.status;
}

// side stream starts after everything queued so far on the main stream
void fork_side(gi_context* ctx) {
    HIP_TRY(hipEventRecord(ctx->ev_fork, ctx->stream));
    HIP_TRY(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
}

// main stream continues after everything queued so far on the side stream
void join_side(gi_context* ctx) {
    HIP_TRY(hipEventRecord(ctx->ev_join, ctx->side));
    HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
}

// timing event on stream s; inside a capture it becomes an event node of the
// graph (external record), so that every replay sets its timestamp
void record(gi_context* ctx, hipEvent_t e, hipStream_t s) {
    HIP_TRY(hipEventRecordWithFlags(e, s, ctx->capturing ? hipEventRecordExternal : hipEventRecordDefault));
}

// the slot's vertices, boxes (and for a new topology keys) must be staged
bool slot_needs_stage(const MeshSlot& m, bool rebuild) {
    return m.set && !m.borrowed && (m.dirty || !m.built || rebuild);
}

// the transform record a slot's stage launch reads from device memory
Xform staged_xform(const MeshSlot& m) {
    Xform x = m.xf;
    for (int k = 0; k < 3; ++k) {
        x.lo[k] = m.src_lo[k];
        x.hi[k] = m.src_hi[k];
    }
    return x;
}

void sync_all(gi_context* ctx) {
    HIP_TRY(hipStreamSynchronize(ctx->side));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
}

// Device copies from the host copies (reconstruct after gi_clear_device).
void upload_mesh(gi_context* ctx, int slot) {
    MeshSlot& m = ctx->mesh[slot];
    if (!m.host_valid) throw GiError{GI_ERR_INTERNAL, "mesh has neither device nor host copies"};
    double* src = m.src.ensure<double>(3 * (size_t)m.nv);
    uint32_t* tri = m.tri.ensure<uint32_t>(3 * (size_t)m.nt);
    hipStream_t st = ctx->stream;
    if (m.nv) HIP_TRY(hipMemcpyAsync(src, m.h_xyz.data(), sizeof(double) * 3 * m.nv, hipMemcpyHostToDevice, st));
    if (m.nt) HIP_TRY(hipMemcpyAsync(tri, m.h_tri.data(), sizeof(uint32_t) * 3 * m.nt, hipMemcpyHostToDevice, st));
    if (m.forest) {
        auto put = [&](DevBuf& d, const std::vector<uint32_t>& h) {
            uint32_t* p = d.ensure<uint32_t>(h.size());
            if (!h.empty())
                HIP_TRY(hipMemcpyAsync(p, h.data(), 4 * h.size(), hipMemcpyHostToDevice, st));
        };
        put(m.voff, m.h_voff);
        put(m.toff, m.h_toff);
        put(m.vtiles, m.h_vtiles);
        put(m.ttiles, m.h_ttiles);
        put(m.vtile0, m.h_vtile0);
        // the host tables must outlive the asynchronous copies
        HIP_TRY(hipStreamSynchronize(st));
    }
    m.on_device = true;
    m.built = false;
    m.dirty = true;
}

// The slot's source arrays (already on the device) back to its host copies.
void ensure_host_copy(gi_context* ctx, int slot) {
    MeshSlot& m = ctx->mesh[slot];
    if (m.host_valid || !m.set) return;
    sync_all(ctx);
    m.h_xyz.resize(3 * (size_t)m.nv);
    m.h_tri.resize(3 * (size_t)m.nt);
    if (m.nv) HIP_TRY(hipMemcpy(m.h_xyz.data(), m.src.get<double>(), 24 * (size_t)m.nv, hipMemcpyDeviceToHost));
    if (m.nt) HIP_TRY(hipMemcpy(m.h_tri.data(), m.tri.get<uint32_t>(), 12 * (size_t)m.nt, hipMemcpyDeviceToHost));
    m.host_valid = true;
}

void drop_host_copy(MeshSlot& m) {
    std::vector<double>().swap(m.h_xyz);
    std::vector<uint32_t>().swap(m.h_tri);
    m.host_valid = false;
}

// After the source arrays of a slot reached the device (src, tri): bounds,
// grid classification and index check in one launch, read back. For a batch
// (tri_local) the indices are per mesh: they are checked against their mesh
// and made global instead. Throws on an out-of-range index.
void inspect_upload(gi_context* ctx, int slot, bool tri_local) {
    MeshSlot& m = ctx->mesh[slot];
    hipStream_t s = ctx->stream;
    double* scratch = ctx->inspect.ensure<double>(inspect_scratch_doubles() + sizeof(MeshInspect) / sizeof(double) + 1);
    MeshInspect* out = (MeshInspect*)(scratch + inspect_scratch_doubles());
    launch_inspect_mesh(m.src.get<double>(), m.nv, m.tri.get<uint32_t>(), tri_local ? 0u : m.nt, scratch, out, s);
    if (tri_local)
        launch_forest_globalize(m.tri.get<uint32_t>(), m.ttiles.get<uint2>(), (uint32_t)(m.h_ttiles.size() / 2),
                                m.toff.get<uint32_t>(), m.voff.get<uint32_t>(), out, s);
    MeshInspect h;
    HIP_TRY(hipMemcpyAsync(&h, out, sizeof(h), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (h.bad_index) {
        m.set = false;
        throw GiError{GI_ERR_INVALID_ARGUMENT, "triangle index out of range"};
    }
    for (int k = 0; k < 3; ++k) {
        m.src_lo[k] = h.lo[k];
        m.src_hi[k] = h.hi[k];
    }
    m.src_grid = h.off_grid ? 0 : 1;
    m.vsrc_ok = false;  // new source vertices
    m.on_device = true;
    m.built = false;
    m.dirty = true;
}

// the integer source copy the unmoved kernels read, when it applies
const int4* source_ints(const MeshSlot& m) {
    return (!m.moved_copy && m.vsrc_ok && !m.borrowed && !m.forest) ? m.vsrc.get<int4>() : nullptr;
}

// A batch: the same stages over all its meshes at once, one launch each.
// Keys mesh << 3 axis bits | Morton: one sort groups by mesh, one Karras build
// makes every mesh's BVH a subtree, whose roots are then looked up.
void stage_forest(gi_context* ctx, int slot) {
    MeshSlot& m = ctx->mesh[slot];
    hipStream_t s = ctx->build_stream[slot];
    uint32_t* grid = ctx->counters.get<uint32_t>() + CNT_COUNT;
    size_t nv = m.nv, nt = m.nt;
    const uint32_t M = m.forest;
    const uint32_t n_vtiles = (uint32_t)(m.h_vtiles.size() / 2), n_ttiles = (uint32_t)(m.h_ttiles.size() / 2);
    double4* v64 = m.v64.ensure<double4>(nv);
    int4* vi = m.vi.ensure<int4>(nv);
    float4* tbox = m.tbox.ensure<float4>(2 * nt);
    m.nodes.ensure<float4>(2 * (nt ? 2 * nt - 1 : 1));
    m.range.ensure<uint2>(nt);
    m.blocks.ensure<float4>(block_levels_elems((uint32_t)nt));
    m.roots.ensure<uint2>(M);
    uint32_t* keys = m.keys.ensure<uint32_t>(nt);
    uint32_t* vals = m.vals.ensure<uint32_t>(nt);
    uint32_t* tk = m.tkeys.ensure<uint32_t>(nt);
    uint32_t* tv = m.tvals.ensure<uint32_t>(nt);
    uint32_t* bounds = m.fbounds.ensure<uint32_t>(6 * (size_t)M);
    float* partials = ctx->bpart[slot].ensure<float>(6 * (size_t)std::max(n_vtiles, kPrepareMaxBlocks));
    launch_prepare_forest(m.src.get<double>(), m.vtiles.get<uint2>(), n_vtiles, m.voff.get<uint32_t>(),
                          m.vtile0.get<uint32_t>(), M, v64, vi, grid + slot, partials, bounds, s);
    uint32_t largest = 0;
    for (uint32_t i = 0; i < M; ++i) largest = std::max(largest, m.h_toff[i + 1] - m.h_toff[i]);
    // mesh id and Morton code within 32 bits (32-bit keys; four 8-bit sort
    // passes for 4,096 meshes, at 6 bits per axis: the batch's meshes are
    // small, and the candidate set does not depend on the tree)
    m.morton_axis = std::min(morton_axis_bits(largest), (32 - bits_for(M)) / 3);  // >= 1: gi_set_batch
    launch_tri_boxes_forest(m.tri.get<uint32_t>(), m.ttiles.get<uint2>(), n_ttiles, m.toff.get<uint32_t>(), v64, vi,
                            grid + slot, bounds, m.morton_axis, tbox, keys, vals, s);
    uint32_t* hist = ctx->bhist[slot].ensure<uint32_t>(sort_hist_elems(nt));
    uint32_t* stmp = ctx->bscan[slot].ensure<uint32_t>(scan_tmp_elems(sort_hist_elems(nt)));
    // the sort runs after the join, together with the other slot's
    ctx->pending_sort[slot] = SortJobU32{keys, vals, tk, tv, m.nt, 3 * m.morton_axis + bits_for(M), false, hist, stmp};
}

// The part of a mesh's build before the sorts: vertices (transform) ->
// boxes -> Morton keys (new topology only). A batch side launches its stages
// on the slot's own stream; a single mesh fills *job for the one staging
// launch both slots share. Returns true when the slot's tree must then be
// built or refit (build_meshes).
constexpr int kQueryMortonBits = 5;

bool stage_mesh(gi_context* ctx, int slot, bool rebuild, StageJob* job) {
    MeshSlot& m = ctx->mesh[slot];
    if (m.borrowed) return false;  // built and kept current by its owner
    if (!m.on_device) upload_mesh(ctx, slot);
    if (!m.dirty && m.built && !rebuild) return false;
    if (m.forest) {
        stage_forest(ctx, slot);
        return true;
    }
    hipStream_t s = ctx->stream;
    uint32_t* grid = ctx->counters.get<uint32_t>() + CNT_COUNT;  // set to 1 by the query's init
    size_t nv = m.nv, nt = m.nt;
    double4* v64 = m.v64.ensure<double4>(nv);
    int4* vi = m.vi.ensure<int4>(nv);
    float4* tbox = m.tbox.ensure<float4>(2 * nt);
    m.nodes.ensure<float4>(2 * (nt ? 2 * nt - 1 : 1));
    m.range.ensure<uint2>(nt);
    m.blocks.ensure<float4>(block_levels_elems((uint32_t)nt));
    uint32_t* keys = m.keys.ensure<uint32_t>(nt);
    uint32_t* vals = m.vals.ensure<uint32_t>(nt);
    uint32_t* tk = m.tkeys.ensure<uint32_t>(nt);
    uint32_t* tv = m.tvals.ensure<uint32_t>(nt);
    const bool topology = !m.built || rebuild;
    Xform* d_xf = ctx->xf_dev.get<Xform>() + slot;
    const Xform x = staged_xform(m);
    if (ctx->capturing || !ctx->xf_current[slot] || memcmp(&x, &ctx->xf_last[slot], sizeof(Xform)) != 0) {
        ctx->h_xf[slot] = x;
        HIP_TRY(hipMemcpyAsync(d_xf, ctx->h_xf + slot, sizeof(Xform), hipMemcpyHostToDevice, s));
        ctx->xf_last[slot] = x;
        ctx->xf_current[slot] = !ctx->capturing;
    }
    // B's triangles in Morton order are only the traversal's queries (no
    // tree): a coarser curve (5 bits per axis, 15-bit keys, two sort passes)
    // keeps a wave's queries together; a tree needs the finer keys
    if (topology) m.morton_axis = (slot == GI_MESH_B && !m.full_tree) ? kQueryMortonBits : morton_axis_bits(m.nt);
    uint32_t* hist = nullptr;
    uint32_t* stmp = nullptr;
    if (topology) {
        hist = ctx->bhist[slot].ensure<uint32_t>(sort_hist_elems(nt));
        stmp = ctx->bscan[slot].ensure<uint32_t>(scan_tmp_elems(sort_hist_elems(nt)));
    }
    // a bucket sort where the key width allows it (its counts come from the
    // staging launch), else the LSD passes (their first counts from there)
    const int bits = 3 * m.morton_axis;
    const bool msd = topology && msd_sort_fits(m.nt, bits);
    const bool hist0 = topology && !msd && radix_sort_takes_hist0(m.nt);
    uint32_t* words = msd ? ctx->msd.get<uint32_t>() + (size_t)slot * kMsdWords : nullptr;
    // an unmoved mesh (no transform, and no snap that would change a vertex)
    // needs no moved copies: no vertex blocks, the grid flag is the upload's
    // classification (set by the query's init), the kernels read m.src
    m.moved_copy = !(x.identity && (!x.snap || m.src_grid == 1) && m.src_grid >= 0);
    if (!m.moved_copy && m.src_grid == 1 && !m.borrowed && !m.forest && !m.vsrc_ok) {
        if (ctx->capturing) throw GiError{GI_ERR_INTERNAL, "integer source copy made during a capture"};
        launch_src_int4(m.src.get<double>(), m.nv, m.vsrc.ensure<int4>(m.nv), s);
        m.vsrc_ok = true;
    }
    *job = StageJob{m.src.get<double>(), m.moved_copy ? m.nv : 0u, m.tri.get<uint32_t>(), m.nt, d_xf, v64, vi,
                    grid + slot,
                    m.morton_axis, tbox, topology ? keys : nullptr, topology ? vals : nullptr,
                    hist0 ? hist : nullptr, !m.moved_copy && m.src_grid == 0 ? 1 : 0, words,
                    msd ? bits - kMsdBits : 0, source_ints(m)};
    // the Morton sort runs after the staging, together with the other slot's
    if (msd)
        ctx->pending_msd[slot] = MsdSortJob{keys, vals, tk, tv, m.nt, bits, words};
    else if (topology)
        ctx->pending_sort[slot] = SortJobU32{keys, vals, tk, tv, m.nt, bits, hist0, hist, stmp};
    return true;
}

template <class K>
BuildJob<K> slot_job(MeshSlot& m) {
    BuildJob<K> j;
    j.keys = m.sorted_in_tmp ? m.tkeys.get<K>() : m.keys.get<K>();
    j.prims = m.sorted_in_tmp ? m.tvals.get<uint32_t>() : m.vals.get<uint32_t>();
    j.n = m.nt;
    j.boxes = m.tbox.get<float4>();
    j.nodes = m.nodes.get<float4>();
    j.range = m.range.get<uint2>();
    j.blocks = m.blocks.get<float4>();
    return j;
}

// Brings both trees up to date. The per-slot stages of the two meshes run
// concurrently (main and side stream); after the join, one launch per build
// kernel serves every tree that needs a build or refit (a refit reruns the
// build on the kept sorted keys: same topology, new boxes). The
// k_build_chunks launch is bracketed by ev_chunk[0..1].
void build_meshes(gi_context* ctx, bool rebuild, int time_mask = 0, bool zero_counters = false) {
    uint32_t* counters = ctx->counters.ensure<uint32_t>(CNT_COUNT + 2);  // shared by both stages: allocate before the fork
    ctx->xf_dev.ensure<Xform>(2);
    ctx->chunk_timed = false;
    ctx->chunk_leaves = 0;
    ctx->stage_timed = false;
    // grid flags of the slots staged below to 1 (their vertices clear them),
    // before the per-slot stages fork
    uint32_t stage_mask = 0;
    for (int sl = 0; sl < 2; ++sl) {
        const MeshSlot& m = ctx->mesh[sl];
        if (!m.forest && slot_needs_stage(m, rebuild)) stage_mask |= 1u << sl;
    }
    // the bucket sorts' totals and cursors start at zero (both slots' words,
    // whichever sort a staged slot takes); the split buckets' words are
    // zeroed once, when allocated (every sort leaves them zero)
    uint32_t* msd_words = nullptr;
    if (stage_mask) {
        const bool fresh = ctx->msd.bytes < sizeof(uint32_t) * 2 * (size_t)kMsdWords;
        if (fresh && ctx->capturing) throw GiError{GI_ERR_INTERNAL, "bucket-sort words allocated during a capture"};
        msd_words = ctx->msd.ensure<uint32_t>(2 * (size_t)kMsdWords);
        if (fresh) HIP_TRY(hipMemsetAsync(msd_words, 0, sizeof(uint32_t) * 2 * (size_t)kMsdWords, ctx->stream));
    }
    if (stage_mask || zero_counters)
        launch_query_init(counters, zero_counters ? CNT_COUNT : 0u, counters + CNT_COUNT, stage_mask, msd_words,
                          msd_words ? 2u : 0u, ctx->stream);
    // batch sides stage on their own streams (fork/join); single meshes share
    // one staging launch on the context's stream
    const bool forest = ctx->mesh[0].forest || ctx->mesh[1].forest;
    if (forest) fork_side(ctx);
    bool need[2] = {false, false};
    StageJob sj[2];
    int nsj = 0;
    for (int sl = 0; sl < 2; ++sl) {
        ctx->pending_sort[sl] = SortJobU32{};
        ctx->pending_msd[sl] = MsdSortJob{};
        sj[nsj].nv = sj[nsj].nt = 0;
        if (ctx->mesh[sl].set) need[sl] = stage_mesh(ctx, sl, rebuild, &sj[nsj]);
        if (need[sl] && !ctx->mesh[sl].forest) ++nsj;
    }
    if (forest) join_side(ctx);
    if (nsj) {
        const bool time_it = (time_mask & GI_TIME_STAGING) != 0;
        if (time_it) record(ctx, ctx->ev_stage[0], ctx->stream);
        launch_stage_meshes(sj, nsj, ctx->stream);
        if (time_it) record(ctx, ctx->ev_stage[1], ctx->stream);
        ctx->stage_timed = time_it;
    }
    {  // both meshes' Morton sorts, pass by pass in shared launches
        SortJobU32 jobs[2];
        int slot_of[2], nj = 0, in_tmp[2] = {0, 0};
        for (int sl = 0; sl < 2; ++sl)
            if (ctx->pending_sort[sl].keys) {
                jobs[nj] = ctx->pending_sort[sl];
                slot_of[nj++] = sl;
            }
        radix_sort_u32_pair(jobs, nj, ctx->stream, in_tmp);
        for (int j = 0; j < nj; ++j) ctx->mesh[slot_of[j]].sorted_in_tmp = in_tmp[j];
        // the bucket sorts, both in the same two launches; they end in (keys, vals)
        MsdSortJob mj[2];
        int nm = 0;
        for (int sl = 0; sl < 2; ++sl)
            if (ctx->pending_msd[sl].keys) {
                mj[nm++] = ctx->pending_msd[sl];
                ctx->mesh[sl].sorted_in_tmp = 0;
            }
        msd_sort_pair(mj, nm, ctx->stream);
    }
    BuildJob<uint32_t> jobs[2];
    int nj = 0;
    for (int sl = 0; sl < 2; ++sl)
        if (need[sl] && (sl == GI_MESH_A || ctx->mesh[sl].full_tree))  // B's queries need its order and boxes only
            jobs[nj++] = slot_job<uint32_t>(ctx->mesh[sl]);
    hipStream_t s = ctx->stream;
    uint32_t leaves = 0;
    for (int j = 0; j < nj; ++j) leaves += jobs[j].n;
    if (nj) {
        const bool time_it = (time_mask & GI_TIME_BVH_CHUNKS) && leaves > 0;
        launch_build_bvh(jobs, nj, (uint32_t)ctx->tune.build_chunk, s, time_it ? ctx->ev_chunk[0] : nullptr,
                         time_it ? ctx->ev_chunk[1] : nullptr,
                         ctx->capturing ? hipEventRecordExternal : hipEventRecordDefault);
        if (time_it) {
            ctx->chunk_timed = true;
            ctx->chunk_leaves = leaves;
        }
    }
    for (int sl = 0; sl < 2; ++sl) {
        if (!need[sl]) continue;
        MeshSlot& m = ctx->mesh[sl];
        if (m.forest && (sl == GI_MESH_A || m.full_tree))  // B's mesh roots are walked by nothing
            launch_forest_roots(m.range.get<uint2>(), m.nodes.get<float4>(), m.nt, m.toff.get<uint32_t>(), m.forest,
                                m.roots.get<uint2>(), s);
        m.built = true;
        m.dirty = false;
    }
    HIP_TRY(hipGetLastError());
}

MeshDev mesh_dev(const MeshSlot& m) {
    MeshDev d;
    d.nv = m.nv;
    d.nt = m.nt;
    d.src = m.src.get<double>();
    d.v64 = m.moved_copy ? m.v64.get<double4>() : nullptr;
    d.vi = m.moved_copy ? m.vi.get<int4>() : source_ints(m);
    d.tri = m.tri.get<uint32_t>();
    d.nodes = m.nodes.get<float4>();
    d.order = m.sorted_in_tmp ? m.tvals.get<uint32_t>() : m.vals.get<uint32_t>();
    d.tbox = m.tbox.get<float4>();
    return d;
}

void ensure_arenas(gi_context* ctx, uint32_t cand, uint32_t exact, uint32_t rec, uint32_t seg) {
    if (cand > ctx->cap_cand) { ctx->cand.ensure<uint2>(cand); ctx->cap_cand = cand; }
    if (exact > ctx->cap_exact) {
        ctx->exact.ensure<uint2>(exact);
        ctx->resid.ensure<uint2>(exact);  // every listed pair may be left to the host
        ctx->slow.ensure<uint2>(exact);
        ctx->cap_exact = exact;
    }
    if (rec > ctx->cap_rec) {
        ctx->rec_key.ensure<uint64_t>(rec);
        ctx->rec_seg.ensure<uint2>(rec);
        ctx->cap_rec = rec;
    }
    if (seg > ctx->cap_seg) {
        ctx->seg_pool.ensure<uint64_t>(kSegPoolWords * (size_t)seg);  // room for the wide form
        ctx->seg_exp.ensure<int32_t>(seg);
        ctx->cap_seg = seg;
    }
}

uint32_t grow(uint64_t need, uint32_t cap) {
    uint64_t v = std::max<uint64_t>(need + need / 4 + 1024, 2 * (uint64_t)cap);
    return (uint32_t)std::min<uint64_t>(v, 0xFFFFFFF0ull);
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}

void empty_result(gi_result* out, int on_device) {
    HostResultStore* st = new HostResultStore();
    st->offs.assign(1, 0);
    out->n_pairs = 0;
    out->n_segments = 0;
    out->pairs = st->pairs.data();
    out->seg_offsets = st->offs.data();
    out->seg_exact = nullptr;
    out->seg_exp2 = nullptr;
    out->seg_f64 = nullptr;
    out->exact_limbs = 2;
    out->on_device = 0;
    out->internal = st;
    (void)on_device;
}

// Everything of a query up to the counter read-back: the trees (when `build`),
// the counter reset, traversal, filter, exact stage and the copy of the
// counters to pinned host memory. `timed` records the stage events (not
// inside a graph capture).
// Segment slots of the output buffers that the pre-read-back output kernels
// write: room for the largest result set they take (at most
// output_small_max() records, each with at most kMaxSegmentsPerPair segments,
// within the arenas).
constexpr uint32_t kMaxSegmentsPerPair = 6;

size_t output_seg_slots(const gi_context* ctx) {
    const uint32_t cap = std::min(output_small_max(), ctx->cap_rec);
    return std::max<size_t>(std::min<size_t>((size_t)kMaxSegmentsPerPair * cap, ctx->cap_seg), 1);
}

// The exact stages a query must launch: only the one the grid flags will
// select when both slots' classification is known on the host.
int exact_kinds(const gi_context* ctx) {
    if (ctx->exact_both) return kExactGrid | kExactWide;
    const int a = predicted_grid(ctx->mesh[0]), b = predicted_grid(ctx->mesh[1]);
    if (a == 1 && b == 1) return kExactGrid;
    if (a == 0 || b == 0) return kExactWide;
    return kExactGrid | kExactWide;
}

// The events a query records (gi_options.stage_timing): every stage, or
// only the launches in the mask.
constexpr int kTimeAll = GI_TIME_STAGING | GI_TIME_BVH_CHUNKS | GI_TIME_TRAVERSE | 16;  // 16: the other stages
int timing_mask(const gi_options& o) { return o.stage_timing == 0 ? kTimeAll : (o.stage_timing & 7); }

// sorted output of a small result set, before the read-back (buffers from
// ensure_output_bufs); its last launch also copies the counters to
// ctx->h_counters
void enqueue_output_small(gi_context* ctx, const gi_options& o) {
    uint32_t* counters = ctx->counters.get<uint32_t>();
    const int bitsB = bits_for(ctx->mesh[1].nt);
    launch_output_small(ctx->rec_key.get<uint64_t>(), ctx->rec_seg.get<uint2>(), counters, ctx->cap_rec, ctx->cap_seg,
                        bits_for(ctx->mesh[0].nt) + bitsB, bitsB, o.want_segments != 0, ctx->pairs.get<uint2>(),
                        ctx->order.get<uint32_t>(), ctx->offs.get<uint32_t>(),
                        ctx->segs.get<uint64_t>(), ctx->exp2.get<int32_t>(), ctx->f64.get<double>(),
                        (uint32_t)output_seg_slots(ctx), ctx->seg_pool.get<uint64_t>(), ctx->seg_exp.get<int32_t>(),
                        counters + CNT_COUNT, ctx->d_counters_host, ctx->stream);
}

void enqueue_presync(gi_context* ctx, const gi_options& o, bool build, bool timed_query) {
    hipStream_t s = ctx->stream;
    const int mask = timed_query ? timing_mask(o) : 0;
    const bool timed = mask == kTimeAll, traverse_timed = (mask & GI_TIME_TRAVERSE) != 0;
    MeshSlot& MA = ctx->mesh[0];
    MeshSlot& MB = ctx->mesh[1];
    const uint32_t M = MA.forest;
    uint32_t* counters = ctx->counters.ensure<uint32_t>(CNT_COUNT + 2);
    if (build) {  // the build's first launch also zeroes the counters
        if (timed) record(ctx, ctx->ev[0], s);
        build_meshes(ctx, o.rebuild != 0, mask, true);
        if (timed) record(ctx, ctx->ev[1], s);
    } else {
        launch_query_init(counters, CNT_COUNT, counters + CNT_COUNT, 0u, nullptr, 0u, s);
    }
    uint32_t* grid = counters + CNT_COUNT;
    MeshDev A = mesh_dev(MA), B = mesh_dev(MB);
    int bitsB = bits_for(MB.nt);
    if (traverse_timed) record(ctx, ctx->ev[6], s);
    const uint32_t* keysB = M ? (MB.sorted_in_tmp ? MB.tkeys.get<uint32_t>() : MB.keys.get<uint32_t>()) : nullptr;
    launch_traverse(A, B, keysB, 3 * MB.morton_axis, M ? MA.roots.get<uint2>() : nullptr, ctx->cand.get<uint2>(),
                    ctx->cap_cand, counters, ctx->tune.traverse_variant, s);
    if (traverse_timed) record(ctx, ctx->ev[2], s);
    launch_filter(A, B, ctx->cand.get<uint2>(), ctx->cap_cand, grid, ctx->exact.get<uint2>(), ctx->cap_exact, counters,
                  ctx->tune.filter_blocks, s);
    if (timed) record(ctx, ctx->ev[3], s);
    ctx->exact_kinds = exact_kinds(ctx);
    // the pairs the one-lane pass leaves (coplanar, degenerate) run after
    // the read-back, only when there are any (run_intersect)
    launch_exact(A, B, ctx->exact.get<uint2>(), ctx->cap_exact, grid, bitsB, ctx->rec_key.get<uint64_t>(),
                 ctx->rec_seg.get<uint2>(), ctx->cap_rec, ctx->seg_pool.get<uint64_t>(), ctx->seg_exp.get<int32_t>(),
                 ctx->cap_seg, ctx->cand.get<uint2>(), ctx->cap_cand, ctx->slow.get<uint2>(), ctx->cap_exact,
                 ctx->resid.get<uint2>(), ctx->cap_exact, counters, ctx->tune.exact_blocks,
                 ctx->exact_kinds | kExactGridSlowLater, s);
    if (timed) record(ctx, ctx->ev[4], s);
    enqueue_output_small(ctx, o);
    if (timed) record(ctx, ctx->ev[5], s);
    HIP_TRY(hipGetLastError());
}

// After a read-back that found grid pairs left by the one-lane exact pass:
// their passes, then the small output again over all records (it rewrites
// the sorted output and the read-back counters). Returns once they are done.
void run_exact_grid_slow(gi_context* ctx, const gi_options& o, bool timed) {
    hipStream_t s = ctx->stream;
    uint32_t* counters = ctx->counters.get<uint32_t>();
    const MeshDev A = mesh_dev(ctx->mesh[0]), B = mesh_dev(ctx->mesh[1]);
    if (timed) record(ctx, ctx->ev[8], s);
    const uint32_t* hgrid = ctx->h_counters + CNT_COUNT;
    if (hgrid[0] && hgrid[1])
        launch_exact_grid_slow(A, B, counters + CNT_COUNT, bits_for(ctx->mesh[1].nt), ctx->rec_key.get<uint64_t>(),
                               ctx->rec_seg.get<uint2>(), ctx->cap_rec, ctx->seg_pool.get<uint64_t>(), ctx->cap_seg,
                               ctx->cand.get<uint2>(), ctx->cap_cand, ctx->slow.get<uint2>(), ctx->cap_exact, counters,
                               ctx->tune.exact_blocks, s);
    else
        launch_exact_wide_slow(A, B, counters + CNT_COUNT, bits_for(ctx->mesh[1].nt), ctx->rec_key.get<uint64_t>(),
                               ctx->rec_seg.get<uint2>(), ctx->cap_rec, ctx->seg_pool.get<uint64_t>(),
                               ctx->seg_exp.get<int32_t>(), ctx->cap_seg, ctx->cand.get<uint2>(), ctx->cap_cand,
                               ctx->slow.get<uint2>(), ctx->cap_exact, ctx->resid.get<uint2>(), ctx->cap_exact,
                               counters, ctx->tune.exact_blocks, s);
    if (timed) record(ctx, ctx->ev[9], s);
    enqueue_output_small(ctx, o);
    if (timed) record(ctx, ctx->ev[5], s);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
}

void ensure_output_bufs(gi_context* ctx, const gi_options& o) {
    const uint32_t cap = std::min(output_small_max(), ctx->cap_rec);
    ctx->pairs.ensure<uint2>(std::max(cap, 1u));
    ctx->order.ensure<uint32_t>(std::max(cap, 1u));
    ctx->offs.ensure<uint32_t>((size_t)cap + 1);
    if (o.want_segments) {
        ctx->segs.ensure<uint64_t>(kSegPoolWords * output_seg_slots(ctx));
        ctx->exp2.ensure<int32_t>(output_seg_slots(ctx));
        ctx->f64.ensure<double>(6 * output_seg_slots(ctx));
    }
}


// Everything enqueue_presync bakes into launches: sizes, flags, device
// pointers (a reallocation changes them), capacities.
std::vector<uint64_t> presync_signature(gi_context* ctx, const gi_options& o) {
    std::vector<uint64_t> v;
    v.push_back((uint64_t)o.rebuild);
    v.push_back((uint64_t)o.want_segments);
    v.push_back((uint64_t)timing_mask(o));
    auto buf = [&](const DevBuf& b) {
        v.push_back((uint64_t)(uintptr_t)b.p);
        v.push_back((uint64_t)b.bytes);
    };
    for (const MeshSlot& m : ctx->mesh) {
        v.push_back(m.nv);
        v.push_back(m.nt);
        v.push_back(m.forest);
        v.push_back(((uint64_t)m.set << 0) | ((uint64_t)m.borrowed << 1) | ((uint64_t)m.on_device << 2) |
                    ((uint64_t)m.built << 3) | ((uint64_t)m.dirty << 4) | ((uint64_t)m.sorted_in_tmp << 5) |
                    ((uint64_t)m.moved_copy << 6) | ((uint64_t)(m.src_grid + 1) << 7));
        v.push_back((uint64_t)m.morton_axis);
        v.push_back(m.generation);
        v.push_back(m.h_vtiles.size());
        v.push_back(m.h_ttiles.size());
        for (const DevBuf* b : {&m.src, &m.tri, &m.v64, &m.vi, &m.tbox, &m.keys, &m.vals, &m.tkeys, &m.tvals,
                                &m.nodes, &m.range, &m.blocks, &m.voff, &m.toff, &m.vtiles, &m.ttiles, &m.vtile0,
                                &m.fbounds, &m.roots})
            buf(*b);
    }
    for (const DevBuf* b : {&ctx->counters, &ctx->xf_dev, &ctx->cand, &ctx->exact, &ctx->slow, &ctx->rec_key,
                            &ctx->rec_seg, &ctx->seg_pool, &ctx->seg_exp, &ctx->resid, &ctx->exp2, &ctx->bhist[0], &ctx->bhist[1], &ctx->bscan[0],
                            &ctx->bscan[1], &ctx->bpart[0], &ctx->bpart[1], &ctx->pairs, &ctx->order, &ctx->offs,
                            &ctx->segs, &ctx->f64, &ctx->tkeys})
        buf(*b);
    v.push_back((uint64_t)ctx->tune.traverse_variant);
    v.push_back((uint64_t)ctx->tune.filter_blocks);
    v.push_back((uint64_t)ctx->tune.exact_blocks);
    v.push_back((uint64_t)ctx->tune.build_chunk);
    v.push_back((uint64_t)exact_kinds(ctx));
    v.push_back(ctx->cap_cand);
    v.push_back(ctx->cap_exact);
    v.push_back(ctx->cap_rec);
    v.push_back(ctx->cap_seg);
    return v;
}

// With use_graph: the first query of a given signature runs eagerly, the
// second is captured into a graph (launched at once), later ones replay it.
// A replay first stages the current transforms in pinned memory (the graph
// copies them to the device) and applies the host-side effects the capture
// had (trees marked current). Returns true when the graph was launched.
bool launch_presync_graph(gi_context* ctx, const gi_options& o) {
    if (!o.use_graph || ctx->graph_off) return false;
    for (const MeshSlot& m : ctx->mesh)
        if (m.set && !m.on_device && !m.borrowed) return false;  // uploads stay outside graphs
    hipStream_t s = ctx->stream;
    std::vector<uint64_t> sig = presync_signature(ctx, o);
    bool need[2];
    for (int sl = 0; sl < 2; ++sl) need[sl] = slot_needs_stage(ctx->mesh[sl], o.rebuild != 0);
    if (ctx->graph_exec && sig == ctx->graph_sig) {
        for (int sl = 0; sl < 2; ++sl)
            if (need[sl]) {
                ctx->h_xf[sl] = staged_xform(ctx->mesh[sl]);
                ctx->xf_current[sl] = false;  // the replay copies h_xf itself
            }
        HIP_TRY(hipGraphLaunch(ctx->graph_exec, s));  // records the stage events itself
        for (int sl = 0; sl < 2; ++sl)
            if (need[sl]) {
                ctx->mesh[sl].built = true;
                ctx->mesh[sl].dirty = false;
            }
        ctx->chunk_timed = ctx->graph_chunk_timed;
        ctx->chunk_leaves = ctx->graph_chunk_leaves;
        ctx->stage_timed = ctx->graph_stage_timed;
        return true;
    }
    if (sig != ctx->graph_sig_seen) {
        ctx->graph_sig_seen = sig;
        return false;
    }
    ctx->drop_graph();
    bool was_built[2], was_dirty[2];
    for (int sl = 0; sl < 2; ++sl) {
        was_built[sl] = ctx->mesh[sl].built;
        was_dirty[sl] = ctx->mesh[sl].dirty;
    }
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        ctx->graph_off = true;
        return false;
    }
    bool ok = true;
    ctx->capturing = true;
    try {
        enqueue_presync(ctx, o, true, true);
    } catch (const GiError&) {
        ok = false;  // the caller runs the same launches eagerly; a real launch error surfaces there
    }
    ctx->capturing = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (!ok || e != hipSuccess || !g ||
        hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0) != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        ctx->graph_exec = nullptr;
        (void)hipGetLastError();
        ctx->graph_off = true;
        // the capture marked the trees current without running anything
        for (int sl = 0; sl < 2; ++sl) {
            ctx->mesh[sl].built = was_built[sl];
            ctx->mesh[sl].dirty = was_dirty[sl];
        }
        return false;
    }
    ctx->graph = g;
    ctx->graph_sig = sig;
    ctx->graph_chunk_timed = ctx->chunk_timed;
    ctx->graph_chunk_leaves = ctx->chunk_leaves;
    ctx->graph_stage_timed = ctx->stage_timed;
    HIP_TRY(hipGraphLaunch(ctx->graph_exec, s));
    return true;
}

// Largest count a 32-bit candidate list can hold (capacities grow up to it).
constexpr uint32_t kMaxListEntries = 0xFFFFFFF0u;

// Decimal-free rendering of one host-stage number: little-endian two's
// complement words as a double.
double words_to_double(const uint64_t* x, int limbs) {
    const bool neg = (x[limbs - 1] >> 63) != 0;
    std::vector<uint64_t> m(x, x + limbs);
    if (neg) {
        uint64_t c = 1;
        for (int l = 0; l < limbs; ++l) {
            m[l] = ~m[l] + c;
            c = (c && m[l] == 0) ? 1 : 0;
        }
    }
    double v = 0.0;
    for (int l = limbs - 1; l >= 0; --l) v = v * 18446744073709551616.0 + (double)m[l];
    return neg ? -v : v;
}

// The end of a query whose exact stage left Q pairs to the host: the GPU's
// sorted result (R pairs, S segments of L limbs, global indices) comes to the
// host, the residual pairs are decided and constructed there
// (host_exact_batch), and the two sorted lists are merged by (a, b). A batch
// is split per mesh pair afterwards.
int finish_with_residual(gi_context* ctx, const gi_options& o, gi_result* out, uint32_t Q, int L, uint32_t R,
                         uint32_t S, const uint2* pairs, const uint32_t* offs, const uint64_t* segs,
                         const int32_t* exp2, const double* f64) {
    hipStream_t s = ctx->stream;
    MeshSlot& MA = ctx->mesh[0];
    MeshSlot& MB = ctx->mesh[1];
    const uint32_t M = MA.forest;
    gi_stats& st = out->stats;
    const bool all_timed = timing_mask(o) == kTimeAll;
    if (all_timed) record(ctx, ctx->ev[5], s);
    const auto t0 = std::chrono::steady_clock::now();
    const size_t words = 8 * (size_t)L;
    std::vector<uint32_t> gp(2 * (size_t)R), go((size_t)R + 1), list(2 * (size_t)Q);
    std::vector<uint64_t> gs(words * S);
    std::vector<int32_t> ge(S, 0);
    std::vector<double> vA(4 * (size_t)MA.nv), vB(4 * (size_t)MB.nv), rawA, rawB;
    if (R) HIP_TRY(hipMemcpyAsync(gp.data(), pairs, 8 * (size_t)R, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(go.data(), offs, 4 * ((size_t)R + 1), hipMemcpyDeviceToHost, s));
    if (S) {
        HIP_TRY(hipMemcpyAsync(gs.data(), segs, 8 * words * S, hipMemcpyDeviceToHost, s));
        if (!st.grid) HIP_TRY(hipMemcpyAsync(ge.data(), exp2, 4 * (size_t)S, hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipMemcpyAsync(list.data(), ctx->resid.get<uint2>(), 8 * (size_t)Q, hipMemcpyDeviceToHost, s));
    // the vertices as the exact stage saw them: the moved copies, or the
    // source of an unmoved mesh (widened to the copies' 4-double layout)
    for (int k = 0; k < 2; ++k) {
        const MeshSlot& mk = ctx->mesh[k];
        std::vector<double>& v = k ? vB : vA;
        std::vector<double>& raw = k ? rawB : rawA;
        if (mk.moved_copy) {
            HIP_TRY(hipMemcpyAsync(v.data(), mk.v64.get<double4>(), 32 * (size_t)mk.nv, hipMemcpyDeviceToHost, s));
        } else {
            raw.resize(3 * (size_t)mk.nv);
            HIP_TRY(hipMemcpyAsync(raw.data(), mk.src.get<double>(), 24 * (size_t)mk.nv, hipMemcpyDeviceToHost, s));
        }
    }
    HIP_TRY(hipStreamSynchronize(s));
    for (int k = 0; k < 2; ++k) {
        const std::vector<double>& raw = k ? rawB : rawA;
        std::vector<double>& v = k ? vB : vA;
        for (size_t i = 0; 3 * i < raw.size(); ++i)
            for (int c = 0; c < 3; ++c) v[4 * i + c] = raw[3 * i + c];
    }
    ensure_host_copy(ctx, 0);
    ensure_host_copy(ctx, 1);
    std::vector<HostHit> hits;
    uint64_t yes = 0, hpred = 0;
    const int nth = ctx->host_threads > 0 ? ctx->host_threads : (int)std::max(1u, std::thread::hardware_concurrency());
    const int hst = host_exact_batch(vA.data(), MA.h_tri.data(), vB.data(), MB.h_tri.data(), list.data(), Q, nth, hits,
                                     &yes, &hpred);
    if (hst == kHostNonFinite) throw GiError{GI_ERR_INVALID_ARGUMENT, "a moved vertex coordinate is not finite"};
    if (hst != 0) throw GiError{GI_ERR_INTERNAL, "FP64 filter certified a pair the exact stage rejects"};
    int LL = L;
    for (const HostHit& h : hits) LL = std::max(LL, h.limbs);
    // merge the two (a, b)-sorted lists
    HostResultStore* hs = new HostResultStore();
    const size_t n = (size_t)R + hits.size();
    hs->pairs.resize(2 * n);
    hs->offs.resize(n + 1);
    size_t nseg_all = S;
    if (o.want_segments)
        for (const HostHit& h : hits) nseg_all += (size_t)h.nseg;
    else
        nseg_all = 0;
    hs->seg.assign(nseg_all * 8 * (size_t)LL, 0);
    hs->exp2.assign(nseg_all, 0);
    hs->f64.resize(6 * nseg_all);
    auto put_number = [&](size_t si, int e, const uint64_t* src, int limbs) {
        uint64_t* dst = &hs->seg[(si * 8 + e) * (size_t)LL];
        const uint64_t ext = (src[limbs - 1] >> 63) ? ~0ull : 0ull;
        for (int l = 0; l < LL; ++l) dst[l] = l < limbs ? src[l] : ext;
    };
    auto render = [&](size_t si) {
        for (int e = 0; e < 2; ++e) {
            const double w = words_to_double(&hs->seg[(si * 8 + 4 * e + 3) * (size_t)LL], LL);
            for (int k = 0; k < 3; ++k)
                hs->f64[6 * si + 3 * e + k] =
                    ldexp(words_to_double(&hs->seg[(si * 8 + 4 * e + k) * (size_t)LL], LL) / w, hs->exp2[si]);
        }
    };
    size_t gi_ = 0, hi = 0, si = 0;
    for (size_t k = 0; k < n; ++k) {
        const bool take_gpu = hi >= hits.size() ||
                              (gi_ < R && ((uint64_t)gp[2 * gi_] << 32 | gp[2 * gi_ + 1]) <
                                              ((uint64_t)hits[hi].a << 32 | hits[hi].b));
        hs->offs[k] = (uint32_t)si;
        if (take_gpu) {
            hs->pairs[2 * k] = gp[2 * gi_];
            hs->pairs[2 * k + 1] = gp[2 * gi_ + 1];
            if (o.want_segments)
                for (uint32_t sg = go[gi_]; sg < go[gi_ + 1]; ++sg, ++si) {
                    for (int e = 0; e < 8; ++e) put_number(si, e, &gs[(sg * 8 + (size_t)e) * L], L);
                    hs->exp2[si] = ge[sg];
                    render(si);
                }
            ++gi_;
        } else {
            const HostHit& h = hits[hi];
            hs->pairs[2 * k] = h.a;
            hs->pairs[2 * k + 1] = h.b;
            st.coplanar += h.coplanar;
            st.degenerate += h.degenerate;
            if (o.want_segments)
                for (int sg = 0; sg < h.nseg; ++sg, ++si) {
                    for (int e = 0; e < 8; ++e) put_number(si, e, &h.data[((size_t)sg * 8 + e) * h.limbs], h.limbs);
                    hs->exp2[si] = h.exp2;
                    render(si);
                }
            ++hi;
        }
    }
    hs->offs[n] = (uint32_t)si;
    if (M) {  // batch: per mesh pair offsets, per-mesh indices
        size_t m = 0;
        ctx->h_pair_off[0] = 0;
        for (size_t i = 0; i < n; ++i) {
            while (m + 1 < M && MA.h_toff[m + 1] <= hs->pairs[2 * i]) ctx->h_pair_off[++m] = (uint32_t)i;
            hs->pairs[2 * i] -= MA.h_toff[m];
            hs->pairs[2 * i + 1] -= MB.h_toff[m];
        }
        while (m + 1 <= M) ctx->h_pair_off[++m] = (uint32_t)n;
    }
    st.host_pairs = Q;
    st.host_yes = yes;
    st.host_predicates = hpred;
    st.pairs = n;
    st.segments = nseg_all;
    st.ms_host = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (all_timed) {
        st.ms_output = elapsed(ctx->ev[4], ctx->ev[5]);  // general doubles: no grid passes after the read-back
        st.ms_total = elapsed(ctx->ev[0], ctx->ev[5]) + st.ms_host;
    }
    out->n_pairs = n;
    out->n_segments = nseg_all;
    out->exact_limbs = LL;
    out->pairs = hs->pairs.data();
    out->seg_offsets = hs->offs.data();
    out->seg_exact = hs->seg.data();
    out->seg_exp2 = hs->exp2.data();
    out->seg_f64 = hs->f64.data();
    out->on_device = 0;
    out->internal = hs;
    if (o.output_on_device) {
        // into the context's device result buffers
        uint32_t* dp = ctx->pairs.ensure<uint32_t>(hs->pairs.size());
        uint32_t* doffs = ctx->offs.ensure<uint32_t>(hs->offs.size());
        uint64_t* dseg = ctx->segs.ensure<uint64_t>(hs->seg.size());
        int32_t* dexp = ctx->exp2.ensure<int32_t>(hs->exp2.size());
        double* df = ctx->f64.ensure<double>(hs->f64.size());
        if (!hs->pairs.empty()) HIP_TRY(hipMemcpy(dp, hs->pairs.data(), 4 * hs->pairs.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(doffs, hs->offs.data(), 4 * hs->offs.size(), hipMemcpyHostToDevice));
        if (nseg_all) {
            HIP_TRY(hipMemcpy(dseg, hs->seg.data(), 8 * hs->seg.size(), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(dexp, hs->exp2.data(), 4 * hs->exp2.size(), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(df, hs->f64.data(), 8 * hs->f64.size(), hipMemcpyHostToDevice));
        }
        delete hs;
        out->pairs = dp;
        out->seg_offsets = doffs;
        out->seg_exact = dseg;
        out->seg_exp2 = dexp;
        out->seg_f64 = df;
        out->on_device = 1;
        out->internal = nullptr;
    }
    return GI_OK;
}

int run_intersect(gi_context* ctx, const gi_options* opt, gi_result* out) {
    memset(out, 0, sizeof(*out));
    gi_options o = {0, 1, 0, 0, 0, 0};
    if (opt) o = *opt;
    MeshSlot& MA = ctx->mesh[0];
    MeshSlot& MB = ctx->mesh[1];
    if (!MA.set || !MB.set) throw GiError{GI_ERR_NOT_READY, "both meshes must be set before gi_intersect"};
    const uint32_t M = MA.forest;
    if (MB.forest != M) throw GiError{GI_ERR_INVALID_ARGUMENT, "mesh slots hold different batches"};
    if (M) ctx->h_pair_off.assign((size_t)M + 1, 0u);
    if (MA.nt == 0 || MB.nt == 0) {  // reference: empty mesh -> early return, no pairs
        empty_result(out, o.output_on_device);
        return GI_OK;
    }
    hipStream_t s = ctx->stream;
    gi_stats& st = out->stats;
    if (ctx->cap_cand == 0) {
        // first guess 8 candidates per triangle, at most 2^26 (the queues
        // grow on overflow; a 2^26 guess is ~3 GB of arenas with segments)
        uint64_t c = ctx->cand_reserve ? ctx->cand_reserve
                                       : std::min<uint64_t>(std::max<uint64_t>(1u << 20, 8ull * (MA.nt + (uint64_t)MB.nt)),
                                                            1ull << 26);
        c = std::min<uint64_t>(c, 0xFFFFFFF0ull);
        ensure_arenas(ctx, (uint32_t)c, (uint32_t)std::max<uint64_t>(c / 4, 1024), (uint32_t)std::max<uint64_t>(c / 4, 1024),
                      (uint32_t)std::max<uint64_t>(c / 4, 1024));
    }
    ctx->counters.ensure<uint32_t>(CNT_COUNT + 2);
    ctx->xf_dev.ensure<Xform>(2);
    int bitsA = bits_for(MA.nt), bitsB = bits_for(MB.nt);
    const uint32_t* hc = ctx->h_counters;
    const uint32_t* hgrid = hc + CNT_COUNT;
    ensure_output_bufs(ctx, o);
    ctx->exact_both = false;
    const bool graphed = launch_presync_graph(ctx, o);
    if (!graphed) enqueue_presync(ctx, o, true, true);
    float slow_ms = 0.f;  // the grid exact stage's passes after the read-back (ms_exact)
    bool slow_ran = false;  // they ran: the output timed from their end (ev[9])
    for (int attempt = 0;; ++attempt) {
        if (attempt) {  // trees are current: queries only
            ensure_output_bufs(ctx, o);
            enqueue_presync(ctx, o, false, true);
        }
        HIP_TRY(hipStreamSynchronize(s));
        if (hc[CNT_OVERFLOW] || hc[CNT_TRAVERSAL_HITS] > kMaxListEntries)
            throw GiError{GI_ERR_TOO_LARGE, "more candidate pairs than a 32-bit list can hold"};
        bool over = hc[CNT_TRAVERSAL_HITS] > ctx->cap_cand || hc[CNT_EXACT_LIST] > ctx->cap_exact ||
                    hc[CNT_RESULTS] > ctx->cap_rec || hc[CNT_SEGMENTS] > ctx->cap_seg;
        // the exact stage the grid flags selected must have run (a host
        // prediction that failed reruns the query with both)
        const bool ran = ctx->exact_kinds & ((hgrid[0] && hgrid[1]) ? kExactGrid : kExactWide);
        if (!ran) ctx->exact_both = true;
        if (!over && ran && hc[CNT_SLOW]) {
            // coplanar and degenerate pairs: their passes and the output
            // again, then the capacities over all records
            run_exact_grid_slow(ctx, o, timing_mask(o) == kTimeAll);
            slow_ms = timing_mask(o) == kTimeAll ? elapsed(ctx->ev[8], ctx->ev[9]) : 0.f;
            slow_ran = true;
            over = hc[CNT_RESULTS] > ctx->cap_rec || hc[CNT_SEGMENTS] > ctx->cap_seg;
        }
        if (!over && ran) break;
        if (attempt > 8) throw GiError{GI_ERR_INTERNAL, "candidate capacity did not converge"};
        uint32_t nc = hc[CNT_TRAVERSAL_HITS] > ctx->cap_cand ? grow(hc[CNT_TRAVERSAL_HITS], ctx->cap_cand) : ctx->cap_cand;
        uint32_t ne = hc[CNT_EXACT_LIST] > ctx->cap_exact ? grow(hc[CNT_EXACT_LIST], ctx->cap_exact) : ctx->cap_exact;
        uint32_t nr = hc[CNT_RESULTS] > ctx->cap_rec ? grow(hc[CNT_RESULTS], ctx->cap_rec) : ctx->cap_rec;
        uint32_t ns = hc[CNT_SEGMENTS] > ctx->cap_seg ? grow(hc[CNT_SEGMENTS], ctx->cap_seg) : ctx->cap_seg;
        ensure_arenas(ctx, nc, ne, nr, ns);
        st.retries++;
    }
    if (hc[CNT_UNSOUND]) throw GiError{GI_ERR_INTERNAL, "FP64 filter certified a pair the exact stage rejects"};
    st.traversal_hits = hc[CNT_TRAVERSAL_HITS];
    st.candidates = hc[CNT_CANDIDATES];
    st.filter_no = hc[CNT_FILTER_NO];
    st.filter_green = hc[CNT_FILTER_GREEN];
    st.filter_yellow = hc[CNT_FILTER_YELLOW];
    st.filter_predicates = hc[CNT_FILTER_PRED];
    st.exact_predicates = hc[CNT_EXACT_PRED];
    st.grid = hgrid[0] && hgrid[1];
    const int tmask = timing_mask(o);
    const bool all_timed = tmask == kTimeAll;
    if (all_timed) st.ms_build = elapsed(ctx->ev[0], ctx->ev[1]);
    if (ctx->chunk_timed) {
        st.ms_bvh_chunks = elapsed(ctx->ev_chunk[0], ctx->ev_chunk[1]);
        st.bvh_chunk_launches = 1;
        st.bvh_chunk_leaves = ctx->chunk_leaves;
    }
    if (ctx->stage_timed) st.ms_stage_meshes = elapsed(ctx->ev_stage[0], ctx->ev_stage[1]);
    if (tmask & GI_TIME_TRAVERSE) st.ms_traverse = elapsed(ctx->ev[6], ctx->ev[2]);
    if (all_timed) st.ms_filter = elapsed(ctx->ev[2], ctx->ev[3]);
    st.graph = graphed ? 1 : 0;

    // the exact stage ran on the GPU: int64/int128 for grid meshes, the wide
    // integer frames for general doubles; pairs that did not fit a frame are
    // the residual, decided on the host and merged below
    const uint32_t Q = hc[CNT_RESIDUAL];
    const int L = st.grid ? kSegLimbsGrid : kSegLimbsWide;
    st.exact_pairs = hc[CNT_EXACT_LIST] - Q;
    st.exact_yellow_yes = hc[CNT_EXACT_YELLOW_YES];
    st.exact_yellow_no = hc[CNT_EXACT_YELLOW_NO];
    st.coplanar = hc[CNT_COPLANAR];
    st.degenerate = hc[CNT_DEGENERATE];
    if (all_timed) st.ms_exact = elapsed(ctx->ev[3], ctx->ev[4]) + slow_ms;
    uint32_t R = hc[CNT_RESULTS], S = o.want_segments ? hc[CNT_SEGMENTS] : 0;
    const size_t words = 8 * (size_t)L;
    // small result sets were sorted and gathered before the read-back
    const bool done = output_small_fits(R, bitsA + bitsB);
    if (!done) {
        uint64_t* keys = ctx->rec_key.get<uint64_t>();
        uint32_t* order = ctx->order.ensure<uint32_t>(R);
        uint64_t* tk = ctx->tkeys.ensure<uint64_t>(R);
        uint32_t* tv = ctx->tvals.ensure<uint32_t>(R);
        uint32_t* hist = ctx->hist.ensure<uint32_t>(sort_hist_elems(R));
        uint32_t* stmp = ctx->scan_tmp.ensure<uint32_t>(scan_tmp_elems(std::max(sort_hist_elems(R), (size_t)R + 1)));
        uint2* pairs = ctx->pairs.ensure<uint2>(R);
        uint32_t* nseg = ctx->nseg.ensure<uint32_t>(R);
        uint32_t* offs = ctx->offs.ensure<uint32_t>((size_t)R + 1);
        uint64_t* segs = ctx->segs.ensure<uint64_t>(words * S);
        int32_t* exp2 = ctx->exp2.ensure<int32_t>(S);
        double* f64 = ctx->f64.ensure<double>(6 * (size_t)S);
        launch_iota(order, R, s);
        if (radix_sort_u64(keys, order, tk, tv, R, bitsA + bitsB, hist, stmp, s)) {
            keys = tk;
            order = tv;
        }
        launch_gather_pairs(keys, order, ctx->rec_seg.get<uint2>(), R, bitsB, pairs, nseg, s);
        if (!o.want_segments) launch_fill_u32(nseg, 0u, R, s);
        exclusive_scan_u32(nseg, offs, R, stmp, offs + R, s);
        if (o.want_segments)
            launch_gather_segments(order, ctx->rec_seg.get<uint2>(), offs, R, ctx->seg_pool.get<uint64_t>(),
                                   ctx->seg_exp.get<int32_t>(), L, segs, exp2, f64, s);
    }
    if (done && S > output_seg_slots(ctx)) {  // more segments than the slots the device gather had
        uint64_t* segs = ctx->segs.ensure<uint64_t>(words * S);
        int32_t* exp2 = ctx->exp2.ensure<int32_t>(S);
        double* f64 = ctx->f64.ensure<double>(6 * (size_t)S);
        launch_gather_segments(ctx->order.get<uint32_t>(), ctx->rec_seg.get<uint2>(), ctx->offs.get<uint32_t>(), R,
                               ctx->seg_pool.get<uint64_t>(), ctx->seg_exp.get<int32_t>(), L, segs, exp2, f64, s);
    }
    uint2* pairs = ctx->pairs.get<uint2>();
    uint32_t* offs = ctx->offs.get<uint32_t>();
    uint64_t* segs = ctx->segs.get<uint64_t>();
    int32_t* exp2 = ctx->exp2.get<int32_t>();
    double* f64 = ctx->f64.get<double>();
    if (Q) return finish_with_residual(ctx, o, out, Q, L, R, S, pairs, offs, segs, exp2, f64);
    const bool host_work = !done || S > output_seg_slots(ctx) || M;
    if (M) {
        uint32_t* poff = ctx->pair_off.ensure<uint32_t>((size_t)M + 1);
        launch_forest_localize(pairs, R, MA.toff.get<uint32_t>(), MB.toff.get<uint32_t>(), M, poff, s);
        HIP_TRY(hipMemcpyAsync(ctx->h_pair_off.data(), poff, 4 * ((size_t)M + 1), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));  // pageable destination
    }
    if (host_work && all_timed) record(ctx, ctx->ev[5], s);
    HIP_TRY(hipGetLastError());
    st.pairs = R;
    st.segments = S;
    out->n_pairs = R;
    out->n_segments = S;
    out->exact_limbs = L;
    if (o.output_on_device) {
        HIP_TRY(hipStreamSynchronize(s));
        out->on_device = 1;
        out->pairs = (uint32_t*)pairs;
        out->seg_offsets = offs;
        out->seg_exact = segs;
        out->seg_exp2 = st.grid ? nullptr : exp2;
        out->seg_f64 = f64;
        out->internal = nullptr;
    } else {
        HostResultStore* hs = new HostResultStore();
        hs->pairs.resize(2 * (size_t)R);
        hs->offs.resize((size_t)R + 1);
        hs->seg.resize(words * S);
        hs->f64.resize(6 * (size_t)S);
        if (!st.grid) hs->exp2.resize(S);
        if (R) HIP_TRY(hipMemcpyAsync(hs->pairs.data(), pairs, 8 * (size_t)R, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(hs->offs.data(), offs, 4 * ((size_t)R + 1), hipMemcpyDeviceToHost, s));
        if (S) {
            HIP_TRY(hipMemcpyAsync(hs->seg.data(), segs, 8 * words * S, hipMemcpyDeviceToHost, s));
            HIP_TRY(hipMemcpyAsync(hs->f64.data(), f64, 48 * (size_t)S, hipMemcpyDeviceToHost, s));
            if (!st.grid) HIP_TRY(hipMemcpyAsync(hs->exp2.data(), exp2, 4 * (size_t)S, hipMemcpyDeviceToHost, s));
        }
        HIP_TRY(hipStreamSynchronize(s));
        out->on_device = 0;
        out->pairs = hs->pairs.data();
        out->seg_offsets = hs->offs.data();
        out->seg_exact = hs->seg.data();
        out->seg_exp2 = st.grid ? nullptr : hs->exp2.data();
        out->seg_f64 = hs->f64.data();
        out->internal = hs;
    }
    if (all_timed) {
        st.ms_output = elapsed(slow_ran ? ctx->ev[9] : ctx->ev[4], ctx->ev[5]);
        st.ms_total = elapsed(ctx->ev[0], ctx->ev[5]);
    }
    return GI_OK;
}

// R = Rz Ry Rx from Euler angles in degrees and T = c + t - R c, evaluated
// in the order of src/CPU/RotationTools.h:28-104 (createRigidTransformation).
Xform rigid_xform(const double rot[3], const double trans[3], const double c[3]) {
    Xform x = identity_xform();
    if (rot[0] == 0.0 && rot[1] == 0.0 && rot[2] == 0.0 && trans[0] == 0.0 && trans[1] == 0.0 && trans[2] == 0.0)
        return x;
    const double DEG_TO_RAD = 3.14159265358979323846 / 180.0;
    double rx = rot[0] * DEG_TO_RAD, ry = rot[1] * DEG_TO_RAD, rz = rot[2] * DEG_TO_RAD;
    double cx = cos(rx), sx = sin(rx), cy = cos(ry), sy = sin(ry), cz = cos(rz), sz = sin(rz);
    double R[3][3];
    R[0][0] = cy * cz;
    R[0][1] = sx * sy * cz - cx * sz;
    R[0][2] = cx * sy * cz + sx * sz;
    R[1][0] = cy * sz;
    R[1][1] = sx * sy * sz + cx * cz;
    R[1][2] = cx * sy * sz - sx * cz;
    R[2][0] = -sy;
    R[2][1] = sx * cy;
    R[2][2] = cx * cy;
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 3; ++k) x.m[r][k] = R[r][k];
        x.m[r][3] = c[r] + trans[r] - (R[r][0] * c[0] + R[r][1] * c[1] + R[r][2] * c[2]);
    }
    x.identity = 0;
    return x;
}

}  // namespace

// ===================================================================== C ABI

extern "C" {

const char* gi_status_string(int status) {
    switch (status) {
        case GI_OK: return "ok";
        case GI_ERR_INVALID_ARGUMENT: return "invalid argument";
        case GI_ERR_HIP: return "HIP runtime error";
        case GI_ERR_OUT_OF_MEMORY: return "out of device memory";
        case GI_ERR_NO_DEVICE: return "no HIP device";
        case GI_ERR_NOT_READY: return "mesh not set";
        case GI_ERR_INTERNAL: return "internal consistency check failed";
        case GI_ERR_TOO_LARGE: return "mesh too large";
        case GI_ERR_IO: return "file cannot be read, parsed or written";
        default: return "unknown status";
    }
}

int gi_create(const gi_config* cfg, gi_context** out) {
    if (!out) return GI_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    gi_context* ctx = new gi_context();
    try {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
            delete ctx;
            return GI_ERR_NO_DEVICE;
        }
        ctx->device = cfg ? cfg->device : 0;
        if (ctx->device < 0 || ctx->device >= ndev) {
            delete ctx;
            return GI_ERR_INVALID_ARGUMENT;
        }
        ctx->host_threads = cfg ? cfg->host_threads : 0;
        ctx->cand_reserve = cfg ? cfg->candidate_reserve : 0;
        HIP_TRY(hipSetDevice(ctx->device));
        HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
        ctx->build_stream[0] = ctx->stream;
        ctx->build_stream[1] = ctx->side;
        HIP_TRY(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
        HIP_TRY(hipHostMalloc((void**)&ctx->h_xf, 2 * sizeof(Xform), hipHostMallocDefault));
        HIP_TRY(hipHostMalloc((void**)&ctx->h_counters, sizeof(uint32_t) * (CNT_COUNT + 2), hipHostMallocDefault));
        memset(ctx->h_counters, 0, sizeof(uint32_t) * (CNT_COUNT + 2));
        HIP_TRY(hipHostGetDevicePointer((void**)&ctx->d_counters_host, ctx->h_counters, 0));
        HIP_TRY(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
        for (auto& e : ctx->ev) HIP_TRY(hipEventCreate(&e));
        for (auto& e : ctx->ev_chunk) HIP_TRY(hipEventCreate(&e));
        for (auto& e : ctx->ev_stage) HIP_TRY(hipEventCreate(&e));
        for (auto& m : ctx->mesh) m.xf = identity_xform();
        *out = ctx;
        return GI_OK;
    } catch (const GiError& e) {
        int st = e.status;
        delete ctx;
        return st;
    }
}

int gi_destroy(gi_context* ctx) {
    if (!ctx) return GI_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->release_device();
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : ctx->ev_chunk)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : ctx->ev_stage)
        if (e) (void)hipEventDestroy(e);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    for (hipStream_t st : ctx->stager.st) (void)hipStreamSynchronize(st);
    for (void* p : ctx->stager.buf) (void)hipHostFree(p);
    for (hipEvent_t e : ctx->stager.ev) (void)hipEventDestroy(e);
    for (hipStream_t st : ctx->stager.st) (void)hipStreamDestroy(st);
    if (ctx->h_xf) (void)hipHostFree(ctx->h_xf);
    if (ctx->h_counters) (void)hipHostFree(ctx->h_counters);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GI_OK;
}

const char* gi_last_error(const gi_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

static int set_mesh_common(gi_context* ctx, int slot, uint32_t nv, uint32_t nt) {
    if (!ctx || slot < 0 || slot > 1) return GI_ERR_INVALID_ARGUMENT;
    if (nt >= (1u << 30)) return GI_ERR_TOO_LARGE;
    (void)nv;
    return GI_OK;
}

int gi_set_mesh(gi_context* ctx, int slot, const double* xyz, uint32_t nv, const uint32_t* tri, uint32_t nt) {
    int st = set_mesh_common(ctx, slot, nv, nt);
    if (st) return st;
    if ((nv && !xyz) || (nt && !tri)) return GI_ERR_INVALID_ARGUMENT;
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        MeshSlot& m = ctx->mesh[slot];
        if (m.borrowed) m.release();
        ++m.generation;
        m.forest = 0;
        drop_host_copy(m);
        m.nv = nv;
        m.nt = nt;
        m.set = true;
        m.xf = identity_xform();
        // straight from the caller's arrays (DMA when they are pinned); the
        // copies complete before the call returns
        double* src = m.src.ensure<double>(3 * (size_t)nv);
        uint32_t* dtri = m.tri.ensure<uint32_t>(3 * (size_t)nt);
        if (nv) HIP_TRY(hipMemcpyAsync(src, xyz, 24 * (size_t)nv, hipMemcpyHostToDevice, ctx->stream));
        if (nt) HIP_TRY(hipMemcpyAsync(dtri, tri, 12 * (size_t)nt, hipMemcpyHostToDevice, ctx->stream));
        inspect_upload(ctx, slot, false);
        return GI_OK;
    } catch (const GiError& e) {
        return fail(ctx, e);
    }
}

int gi_set_mesh_device(gi_context* ctx, int slot, const double* d_xyz, uint32_t nv, const uint32_t* d_tri,
                       uint32_t nt) {
    int st = set_mesh_common(ctx, slot, nv, nt);
    if (st) return st;
    if ((nv && !d_xyz) || (nt && !d_tri)) return GI_ERR_INVALID_ARGUMENT;
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        MeshSlot& m = ctx->mesh[slot];
        if (m.borrowed) m.release();
        ++m.generation;
        m.forest = 0;
        m.nv = nv;
        m.nt = nt;
        m.set = true;
        m.xf = identity_xform();
        drop_host_copy(m);
        double* src = m.src.ensure<double>(3 * (size_t)nv);
        uint32_t* tri = m.tri.ensure<uint32_t>(3 * (size_t)nt);
        if (nv) HIP_TRY(hipMemcpyAsync(src, d_xyz, 24 * (size_t)nv, hipMemcpyDeviceToDevice, ctx->stream));
        if (nt) HIP_TRY(hipMemcpyAsync(tri, d_tri, 12 * (size_t)nt, hipMemcpyDeviceToDevice, ctx->stream));
        inspect_upload(ctx, slot, false);
        return GI_OK;
    } catch (const GiError& e) {
        return fail(ctx, e);
    }
}

int gi_set_transform(gi_context* ctx, int slot, const double rot_deg[3], const double trans[3], const double center[3],
                     int snap_to_grid) {
    if (!ctx || slot < 0 || slot > 1 || !rot_deg || !trans || !center) return GI_ERR_INVALID_ARGUMENT;
    if (!ctx->mesh[slot].set) return GI_ERR_NOT_READY;
    if (ctx->mesh[slot].borrowed) return GI_ERR_INVALID_ARGUMENT;  // a shared mesh is read-only
    if (ctx->mesh[slot].forest) return GI_ERR_INVALID_ARGUMENT;    // batches take placed meshes
    Xform x = rigid_xform(rot_deg, trans, center);
    x.snap = snap_to_grid ? 1 : 0;
    ctx->mesh[slot].xf = x;
    ctx->mesh[slot].dirty = true;
    return GI_OK;
}

int gi_set_transform_matrix(gi_context* ctx, int slot, const double m[12], int snap_to_grid) {
    if (!ctx || slot < 0 || slot > 1 || !m) return GI_ERR_INVALID_ARGUMENT;
    if (!ctx->mesh[slot].set) return GI_ERR_NOT_READY;
    if (ctx->mesh[slot].borrowed) return GI_ERR_INVALID_ARGUMENT;  // a shared mesh is read-only
    if (ctx->mesh[slot].forest) return GI_ERR_INVALID_ARGUMENT;    // batches take placed meshes
    Xform x = identity_xform();
    bool ident = true;
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 4; ++k) {
            x.m[r][k] = m[4 * r + k];
            double want = (k == r) ? 1.0 : 0.0;
            if (m[4 * r + k] != want) ident = false;
        }
    x.identity = ident ? 1 : 0;
    x.snap = snap_to_grid ? 1 : 0;
    ctx->mesh[slot].xf = x;
    ctx->mesh[slot].dirty = true;
    return GI_OK;
}

int gi_build(gi_context* ctx) {
    if (!ctx) return GI_ERR_INVALID_ARGUMENT;
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        build_meshes(ctx, false);
        sync_all(ctx);
        return GI_OK;
    } catch (const GiError& e) {
        return fail(ctx, e);
    }
}

int gi_get_vertices(gi_context* ctx, int slot, double* xyz, uint32_t nv) {
    if (!ctx || slot < 0 || slot > 1 || (nv && !xyz)) return GI_ERR_INVALID_ARGUMENT;
    MeshSlot& m = ctx->mesh[slot];
    if (!m.set) return GI_ERR_NOT_READY;
    if (nv != m.nv) {
        ctx->err = "vertex count differs from the slot's";
        return GI_ERR_INVALID_ARGUMENT;
    }
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        build_meshes(ctx, false);
        sync_all(ctx);
        if (!m.moved_copy) {  // unmoved: the source vertices are the ones the query uses
            if (nv) HIP_TRY(hipMemcpy(xyz, m.src.get<double>(), 24 * (size_t)nv, hipMemcpyDeviceToHost));
            return GI_OK;
        }
        std::vector<double> v4(4 * (size_t)nv);
        if (nv)
            HIP_TRY(hipMemcpy(v4.data(), m.v64.get<double4>(), 32 * (size_t)nv, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < nv; ++i)
            for (int k = 0; k < 3; ++k) xyz[3 * i + k] = v4[4 * i + k];
        return GI_OK;
    } catch (const GiError& e) {
        return fail(ctx, e);
    }
}

int gi_get_tree(gi_context* ctx, int slot, uint32_t n, uint32_t* keys, uint32_t* ids, uint32_t* left,
                uint32_t* rope, uint32_t* range, uint32_t* leaf_rope) {
    if (!ctx || slot < 0 || slot > 1) return GI_ERR_INVALID_ARGUMENT;
    MeshSlot& m = ctx->mesh[slot];
    if (!m.set) return GI_ERR_NOT_READY;
    if (n != m.nt || m.forest) {
        ctx->err = m.forest ? "the slot holds a batch" : "triangle count differs from the slot's";
        return GI_ERR_INVALID_ARGUMENT;
    }
    const bool tree = left || rope || range || leaf_rope;
    if (tree && slot == GI_MESH_B && !m.full_tree) {
        ctx->err = "mesh B builds no tree";
        return GI_ERR_NOT_READY;
    }
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        build_meshes(ctx, false);
        sync_all(ctx);
        if (n == 0) return GI_OK;
        const uint32_t* dk = m.sorted_in_tmp ? m.tkeys.get<uint32_t>() : m.keys.get<uint32_t>();
        const uint32_t* dv = m.sorted_in_tmp ? m.tvals.get<uint32_t>() : m.vals.get<uint32_t>();
        if (keys) HIP_TRY(hipMemcpy(keys, dk, 4 * (size_t)n, hipMemcpyDeviceToHost));
        if (ids) HIP_TRY(hipMemcpy(ids, dv, 4 * (size_t)n, hipMemcpyDeviceToHost));
        if (!tree) return GI_OK;
        std::vector<float4> nodes(2 * (2 * (size_t)n - 1));
        HIP_TRY(hipMemcpy(nodes.data(), m.nodes.get<float4>(), nodes.size() * sizeof(float4), hipMemcpyDeviceToHost));
        auto word = [](float w) {
            uint32_t u;
            memcpy(&u, &w, 4);
            return u;
        };
        for (size_t i = 0; i + 1 < n; ++i) {
            if (left) left[i] = word(nodes[2 * i].w);
            if (rope) rope[i] = word(nodes[2 * i + 1].w);
        }
        if (leaf_rope)
            for (size_t k = 0; k < n; ++k) leaf_rope[k] = word(nodes[2 * (n - 1 + k) + 1].w);
        if (range && n > 1) HIP_TRY(hipMemcpy(range, m.range.get<uint2>(), 8 * (size_t)(n - 1), hipMemcpyDeviceToHost));
        return GI_OK;
    } catch (const GiError& e) {
        return fail(ctx, e);
    }
}

int gi_intersect(gi_context* ctx, const gi_options* opt, gi_result* out) {
    if (!ctx || !out) return GI_ERR_INVALID_ARGUMENT;
    if (ctx->mesh[0].forest || ctx->mesh[1].forest) {
        ctx->err = "the context holds a batch: use gi_intersect_batch";
        return GI_ERR_INVALID_ARGUMENT;
    }
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        return run_intersect(ctx, opt, out);
    } catch (const GiError& e) {
        return fail(ctx, e);
    }
}

int gi_result_free(gi_result* res) {
    if (!res) return GI_ERR_INVALID_ARGUMENT;
    delete (HostResultStore*)res->internal;
    memset(res, 0, sizeof(*res));
    return GI_OK;
}

int gi_intersect_oneshot(const gi_config* cfg, const double* xyzA, uint32_t nvA, const uint32_t* triA, uint32_t ntA,
                         const double* xyzB, uint32_t nvB, const uint32_t* triB, uint32_t ntB, const gi_options* opt,
                         gi_result* out) {
    gi_context* ctx = nullptr;
    int st = gi_create(cfg, &ctx);
    if (st) return st;
    gi_options o = {0, 1, 0, 0, 0, 0};
    if (opt) o = *opt;
    o.output_on_device = 0;  // the context dies with this call
    st = gi_set_mesh(ctx, 0, xyzA, nvA, triA, ntA);
    if (!st) st = gi_set_mesh(ctx, 1, xyzB, nvB, triB, ntB);
    if (!st) st = gi_intersect(ctx, &o, out);
    gi_destroy(ctx);
    return st;
}

int gi_share_mesh(gi_context* dst, int dst_slot, gi_context* src, int src_slot) {
    if (!dst || !src || dst == src || dst_slot < 0 || dst_slot > 1 || src_slot < 0 || src_slot > 1)
        return GI_ERR_INVALID_ARGUMENT;
    if (dst->device != src->device) return GI_ERR_INVALID_ARGUMENT;
    if (!src->mesh[src_slot].set) return GI_ERR_NOT_READY;
    if (src->mesh[src_slot].forest) return GI_ERR_INVALID_ARGUMENT;
    try {
        HIP_TRY(hipSetDevice(src->device));
        MeshSlot& sm = src->mesh[src_slot];
        if (dst_slot == GI_MESH_A && !sm.full_tree && src_slot == GI_MESH_B) {
            if (sm.borrowed) return GI_ERR_INVALID_ARGUMENT;  // its owner built no tree for it
            // an A slot walks the tree: build B's tree from now on, on the
            // finer Morton keys a tree needs (a new topology)
            sm.full_tree = true;
            sm.dirty = true;
            sm.built = false;
        }
        build_meshes(src, false);
        sync_all(src);
        MeshSlot& d = dst->mesh[dst_slot];
        const MeshSlot& s = src->mesh[src_slot];
        d.release();  // bumps the generation
        drop_host_copy(d);  // made from the shared device arrays when needed
        d.nv = s.nv;
        d.nt = s.nt;
        d.set = true;
        d.on_device = true;
        d.built = true;
        d.dirty = false;
        d.xf = s.xf;
        for (int k = 0; k < 3; ++k) { d.src_lo[k] = s.src_lo[k]; d.src_hi[k] = s.src_hi[k]; }
        d.src_grid = s.src_grid;
        d.moved_copy = s.moved_copy;
        d.sorted_in_tmp = s.sorted_in_tmp;
        d.morton_axis = s.morton_axis;
        d.full_tree = s.full_tree;
        d.borrowed = true;
        DevBuf* dd[] = {&d.src, &d.tri, &d.v64, &d.vi, &d.tbox, &d.keys, &d.vals, &d.tkeys, &d.tvals, &d.nodes,
                        &d.range, &d.blocks};
        const DevBuf* ss[] = {&s.src, &s.tri, &s.v64, &s.vi, &s.tbox, &s.keys, &s.vals, &s.tkeys, &s.tvals,
                              &s.nodes, &s.range, &s.blocks};
        static_assert(sizeof(dd) / sizeof(dd[0]) == sizeof(ss) / sizeof(ss[0]), "buffer lists differ");
        for (size_t i = 0; i < sizeof(dd) / sizeof(dd[0]); ++i) { dd[i]->p = ss[i]->p; dd[i]->bytes = ss[i]->bytes; }
        // the owner's grid classification of this mesh
        uint32_t* dgrid = dst->counters.ensure<uint32_t>(CNT_COUNT + 2) + CNT_COUNT;
        const uint32_t* sgrid = src->counters.get<uint32_t>() + CNT_COUNT;
        HIP_TRY(hipSetDevice(dst->device));
        HIP_TRY(hipMemcpyAsync(dgrid + dst_slot, sgrid + src_slot, sizeof(uint32_t), hipMemcpyDeviceToDevice,
                               dst->stream));
        HIP_TRY(hipStreamSynchronize(dst->stream));
        return GI_OK;
    } catch (const GiError& e) {
        return fail(dst, e);
    }
}

namespace {

// One side of a batch -> the slot's sizes and tile tables (the meshes stored
// back to back: mesh i's vertices from h_voff[i], triangles from h_toff[i]).
void fill_forest(MeshSlot& m, const gi_mesh_pair* pairs, uint32_t n, bool side_b) {
    uint64_t nv = 0, nt = 0;
    for (uint32_t i = 0; i < n; ++i) {
        nv += side_b ? pairs[i].nvB : pairs[i].nvA;
        nt += side_b ? pairs[i].ntB : pairs[i].ntA;
    }
    // keep the device buffers (a later batch of the same size reuses them;
    // freeing and mapping tens of GB again costs more than the upload)
    if (m.borrowed) m.release();
    ++m.generation;
    m.full_tree = false;
    drop_host_copy(m);
    m.forest = n;
    m.nv = (uint32_t)nv;
    m.nt = (uint32_t)nt;
    m.h_voff.assign((size_t)n + 1, 0u);
    m.h_toff.assign((size_t)n + 1, 0u);
    m.h_vtiles.clear();
    m.h_ttiles.clear();
    m.h_vtile0.assign((size_t)n + 1, 0u);
    uint32_t vo = 0, to = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t mv = side_b ? pairs[i].nvB : pairs[i].nvA, mt = side_b ? pairs[i].ntB : pairs[i].ntA;
        m.h_vtile0[i] = (uint32_t)(m.h_vtiles.size() / 2);
        for (uint32_t f = 0; f < mv; f += kForestTile) {
            m.h_vtiles.push_back(i);
            m.h_vtiles.push_back(vo + f);
        }
        for (uint32_t f = 0; f < mt; f += kForestTile) {
            m.h_ttiles.push_back(i);
            m.h_ttiles.push_back(to + f);
        }
        vo += mv;
        to += mt;
        m.h_voff[i + 1] = vo;
        m.h_toff[i + 1] = to;
    }
    m.h_vtile0[n] = (uint32_t)(m.h_vtiles.size() / 2);
    m.set = true;
    m.on_device = false;
    m.built = false;
    m.dirty = true;
    m.xf = identity_xform();
}

// Host arrays copied back to back into one device range through pinned
// memory. Each of up to kStageWorkers host threads takes a contiguous share
// of the range and moves it in kStageChunk pieces: its own two pinned
// buffers take turns, so that its memcpy into one overlaps the DMA out of
// the other (on the worker's own stream), and the workers' DMAs overlap each
// other. Copies from pageable memory (hipMemcpyAsync on the caller's array)
// take the runtime's own staging path one copy at a time, which left a
// 6.6 GB batch upload at 20 GB/s. Returns once everything is on the device.
struct HostPiece {
    const void* p;
    size_t bytes;
};
constexpr int kStageWorkers = 8;
constexpr size_t kStageChunk = 16u << 20;

void ensure_stager(gi_context* ctx) {
    auto& S = ctx->stager;
    if (S.workers) return;
    const int hw = ctx->host_threads > 0 ? ctx->host_threads : (int)std::max(1u, std::thread::hardware_concurrency());
    const int w = std::max(1, std::min(kStageWorkers, hw));
    for (int i = 0; i < w; ++i) {
        hipStream_t st;
        HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        S.st.push_back(st);
        for (int k = 0; k < 2; ++k) {
            void* p = nullptr;
            HIP_TRY(hipHostMalloc(&p, kStageChunk, hipHostMallocDefault));
            S.buf.push_back(p);
            hipEvent_t e;
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            S.ev.push_back(e);
        }
    }
    S.workers = w;
}

void upload_gather(gi_context* ctx, char* dst, const std::vector<HostPiece>& pieces) {
    std::vector<size_t> off(pieces.size() + 1, 0);
    for (size_t i = 0; i < pieces.size(); ++i) off[i + 1] = off[i] + pieces[i].bytes;
    const size_t total = off.back();
    if (total == 0) return;
    ensure_stager(ctx);
    auto& S = ctx->stager;
    const int W = (int)std::min<size_t>((size_t)S.workers, (total + kStageChunk - 1) / kStageChunk);
    std::vector<int> status(W, GI_OK);
    auto work = [&](int w) {
        try {
            HIP_TRY(hipSetDevice(ctx->device));
            const size_t lo = total * (size_t)w / W, hi = total * (size_t)(w + 1) / W;
            size_t pi = (size_t)(std::upper_bound(off.begin(), off.end(), lo) - off.begin()) - 1;
            int turn = 0;
            for (size_t a = lo; a < hi; a += kStageChunk, turn ^= 1) {
                const size_t len = std::min(kStageChunk, hi - a);
                char* buf = (char*)S.buf[2 * w + turn];
                hipEvent_t ev = S.ev[2 * w + turn];
                HIP_TRY(hipEventSynchronize(ev));  // this buffer's previous DMA is done
                for (size_t got = 0; got < len;) {
                    while (off[pi + 1] <= a + got) ++pi;
                    const size_t in = a + got - off[pi];
                    const size_t take = std::min(len - got, pieces[pi].bytes - in);
                    memcpy(buf + got, (const char*)pieces[pi].p + in, take);
                    got += take;
                }
                HIP_TRY(hipMemcpyAsync(dst + a, buf, len, hipMemcpyHostToDevice, S.st[w]));
                HIP_TRY(hipEventRecord(ev, S.st[w]));
            }
            HIP_TRY(hipStreamSynchronize(S.st[w]));
        } catch (const GiError& e) {
            status[w] = e.status;
        }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < W; ++w) th.emplace_back(work, w);
    work(0);
    for (auto& t : th) t.join();
    for (int w = 0; w < W; ++w)
        if (status[w] != GI_OK) throw GiError{status[w], "staged upload failed"};
}

// Batches above this many bytes per array go through upload_gather.
constexpr size_t kStagedUploadMin = 16u << 20;

// The batch's arrays straight from the caller's, each mesh to its place in
// the slot's device arrays (indices still per mesh), the tile tables, then
// the device check that makes the indices global.
void upload_forest(gi_context* ctx, int slot, const gi_mesh_pair* pairs, uint32_t n, bool side_b) {
    MeshSlot& m = ctx->mesh[slot];
    hipStream_t st = ctx->stream;
    double* src = m.src.ensure<double>(3 * (size_t)m.nv);
    uint32_t* tri = m.tri.ensure<uint32_t>(3 * (size_t)m.nt);
    if (24 * (size_t)m.nv + 12 * (size_t)m.nt >= kStagedUploadMin) {
        // the device arrays hold the meshes back to back in pair order
        std::vector<HostPiece> xyz, idx;
        for (uint32_t i = 0; i < n; ++i) {
            xyz.push_back({side_b ? pairs[i].xyzB : pairs[i].xyzA, 24 * (size_t)(side_b ? pairs[i].nvB : pairs[i].nvA)});
            idx.push_back({side_b ? pairs[i].triB : pairs[i].triA, 12 * (size_t)(side_b ? pairs[i].ntB : pairs[i].ntA)});
        }
        HIP_TRY(hipStreamSynchronize(st));  // nothing queued may still read the arrays being replaced
        upload_gather(ctx, (char*)src, xyz);
        upload_gather(ctx, (char*)tri, idx);
    } else {
        for (uint32_t i = 0; i < n; ++i) {
            const double* xyz = side_b ? pairs[i].xyzB : pairs[i].xyzA;
            const uint32_t* t = side_b ? pairs[i].triB : pairs[i].triA;
            const uint32_t mv = side_b ? pairs[i].nvB : pairs[i].nvA, mt = side_b ? pairs[i].ntB : pairs[i].ntA;
            if (mv)
                HIP_TRY(hipMemcpyAsync(src + 3 * (size_t)m.h_voff[i], xyz, 24 * (size_t)mv, hipMemcpyHostToDevice, st));
            if (mt)
                HIP_TRY(hipMemcpyAsync(tri + 3 * (size_t)m.h_toff[i], t, 12 * (size_t)mt, hipMemcpyHostToDevice, st));
        }
    }
    auto put = [&](DevBuf& d, const std::vector<uint32_t>& h) {
        uint32_t* p = d.ensure<uint32_t>(h.size());
        if (!h.empty()) HIP_TRY(hipMemcpyAsync(p, h.data(), 4 * h.size(), hipMemcpyHostToDevice, st));
    };
    put(m.voff, m.h_voff);
    put(m.toff, m.h_toff);
    put(m.vtiles, m.h_vtiles);
    put(m.ttiles, m.h_ttiles);
    put(m.vtile0, m.h_vtile0);
    inspect_upload(ctx, slot, true);  // synchronises: the caller's arrays and the tables may go
}

}  // namespace

int gi_set_batch(gi_context* ctx, const gi_mesh_pair* pairs, uint32_t n_pairs) {
    if (!ctx || !pairs || n_pairs == 0) return GI_ERR_INVALID_ARGUMENT;
    uint64_t nv[2] = {0, 0}, nt[2] = {0, 0};
    for (uint32_t i = 0; i < n_pairs; ++i) {
        const gi_mesh_pair& p = pairs[i];
        const double* xyz[2] = {p.xyzA, p.xyzB};
        const uint32_t* tri[2] = {p.triA, p.triB};
        uint32_t mv[2] = {p.nvA, p.nvB}, mt[2] = {p.ntA, p.ntB};
        for (int sd = 0; sd < 2; ++sd) {
            if ((mv[sd] && !xyz[sd]) || (mt[sd] && !tri[sd])) {
                ctx->err = "mesh pair " + std::to_string(i) + ": null array";
                return GI_ERR_INVALID_ARGUMENT;
            }
            nv[sd] += mv[sd];
            nt[sd] += mt[sd];
        }
    }
    // 2^30 triangles per side: node indices and the status bits of the
    // candidate lists; 2^32 vertices: global indices; 2^29 pairs: the mesh id
    // leaves at least one Morton bit per axis in the 32-bit BVH keys
    for (int sd = 0; sd < 2; ++sd)
        if (nt[sd] >= (1ull << 30) || nv[sd] >= (1ull << 32)) return GI_ERR_TOO_LARGE;
    if (n_pairs > (1u << 29)) return GI_ERR_TOO_LARGE;
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        sync_all(ctx);
        for (int sd = 0; sd < 2; ++sd) {
            if (ctx->mesh[sd].borrowed) ctx->mesh[sd].release();
            fill_forest(ctx->mesh[sd], pairs, n_pairs, sd == 1);
            upload_forest(ctx, sd, pairs, n_pairs, sd == 1);
        }
        sync_all(ctx);
        return GI_OK;
    } catch (const GiError& e) {
        for (auto& m : ctx->mesh) m.set = false;
        return fail(ctx, e);
    }
}

int gi_intersect_batch(gi_context* ctx, const gi_options* opt, gi_result* out, uint64_t* pair_offsets) {
    if (!ctx || !out) return GI_ERR_INVALID_ARGUMENT;
    if (!ctx->mesh[0].forest || ctx->mesh[0].forest != ctx->mesh[1].forest) {
        ctx->err = "no batch set (gi_set_batch)";
        return GI_ERR_NOT_READY;
    }
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        int st = run_intersect(ctx, opt, out);
        if (pair_offsets)
            for (size_t i = 0; i < ctx->h_pair_off.size(); ++i) pair_offsets[i] = ctx->h_pair_off[i];
        return st;
    } catch (const GiError& e) {
        return fail(ctx, e);
    }
}

int gi_get_tuning(const gi_context* ctx, gi_tuning* out) {
    if (!ctx || !out) return GI_ERR_INVALID_ARGUMENT;
    *out = ctx->tune;
    return GI_OK;
}

int gi_set_tuning(gi_context* ctx, const gi_tuning* t) {
    if (!ctx || !t) return GI_ERR_INVALID_ARGUMENT;
    if (t->traverse_variant < 0 || t->traverse_variant > 3 || t->filter_blocks < 1 || t->filter_blocks > 65535 ||
        t->exact_blocks < 1 || t->exact_blocks > 65535 || (t->build_chunk != 256 && t->build_chunk != 512)) {
        ctx->err = "tuning out of range";
        return GI_ERR_INVALID_ARGUMENT;
    }
    ctx->tune = *t;
    return GI_OK;
}

// Coordinate search over the launch shapes on the context's current meshes:
// for each knob in turn (build chunk, traversal shape, filter width, exact
// width), every candidate value is timed on `reps` queries by the event
// timing of the stage it affects (median), and the fastest is kept. Results
// do not depend on these settings, only the time does.
int gi_autotune(gi_context* ctx, const gi_options* opt, int32_t reps, gi_tuning* out) {
    if (!ctx) return GI_ERR_INVALID_ARGUMENT;
    if (!ctx->mesh[0].set || !ctx->mesh[1].set) return GI_ERR_NOT_READY;
    if (reps <= 0) reps = 3;
    gi_options o = {1, 1, 0, 0, 0, 0};
    if (opt) o = *opt;
    o.output_on_device = 1;
    o.use_graph = 0;
    o.stage_timing = 0;  // every stage timed: the stage times pick the launch shapes
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        auto query = [&](bool rebuild) {
            gi_result res;
            gi_options q = o;
            q.rebuild = rebuild ? 1 : o.rebuild;
            run_intersect(ctx, &q, &res);
            gi_stats st = res.stats;
            gi_result_free(&res);
            return st;
        };
        auto median_ms = [&](int stage) {
            std::vector<float> v;
            for (int r = 0; r < reps; ++r) {
                gi_stats st = query(stage == 0);
                v.push_back(stage == 0 ? st.ms_build : stage == 1 ? st.ms_traverse : stage == 2 ? st.ms_filter
                                                                                                : st.ms_exact);
            }
            std::sort(v.begin(), v.end());
            return v[v.size() / 2];
        };
        query(true);  // arenas reach their sizes
        auto pick = [&](int32_t* knob, const std::vector<int32_t>& values, int stage) {
            int32_t best = *knob;
            float best_ms = 1e30f;
            for (int32_t val : values) {
                *knob = val;
                float ms = median_ms(stage);
                if (ms < best_ms) {
                    best_ms = ms;
                    best = val;
                }
            }
            *knob = best;
        };
        pick(&ctx->tune.build_chunk, {256, 512}, 0);
        pick(&ctx->tune.traverse_variant, {0, 1, 2, 3}, 1);
        pick(&ctx->tune.filter_blocks, {512, 1024, 2048, 4096, 8192}, 2);
        pick(&ctx->tune.exact_blocks, {256, 512, 1024, 2048, 4096}, 3);
        if (out) *out = ctx->tune;
        return GI_OK;
    } catch (const GiError& e) {
        return fail(ctx, e);
    }
}

int gi_clear_device(gi_context* ctx) {
    if (!ctx) return GI_ERR_INVALID_ARGUMENT;
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        for (int sl = 0; sl < 2; ++sl)  // the host copies reconstruct the device state
            if (!ctx->mesh[sl].borrowed) ensure_host_copy(ctx, sl);
    } catch (const GiError& e) {
        return fail(ctx, e);
    }
    (void)hipStreamSynchronize(ctx->side);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->release_device();
    return GI_OK;
}

int gi_reconstruct_device(gi_context* ctx) {
    if (!ctx) return GI_ERR_INVALID_ARGUMENT;
    for (int sl = 0; sl < 2; ++sl)
        if (!ctx->mesh[sl].set) {
            ctx->err = "reconstruct needs both meshes";  // reference prints to stderr and returns
            return GI_ERR_NOT_READY;
        }
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        build_meshes(ctx, true);
        sync_all(ctx);
        return GI_OK;
    } catch (const GiError& e) {
        return fail(ctx, e);
    }
}

int gi_exact_pair_host(const double a[9], const double b[9], int32_t* hit, int32_t* coplanar, int32_t* nseg,
                       int32_t* exp2, int32_t* limbs, uint64_t* seg, int32_t limbs_cap) {
    if (!a || !b || !hit || !nseg || !exp2 || !limbs) return GI_ERR_INVALID_ARGUMENT;
    HostHit h;
    bool ok;
    bool r = host_exact_pair(a, b, h, &ok);
    if (!ok) return GI_ERR_INVALID_ARGUMENT;  // a non-finite coordinate
    *hit = r ? 1 : 0;
    if (coplanar) *coplanar = r ? h.coplanar : 0;
    *nseg = r ? h.nseg : 0;
    *exp2 = r ? h.exp2 : 0;
    *limbs = r ? h.limbs : 0;
    if (r && seg) {
        if (h.limbs > limbs_cap) return GI_ERR_TOO_LARGE;
        for (int s = 0; s < h.nseg; ++s)
            for (int e = 0; e < 8; ++e) {
                const uint64_t* src = &h.data[((size_t)s * 8 + e) * h.limbs];
                uint64_t* dst = seg + ((size_t)s * 8 + e) * limbs_cap;
                uint64_t ext = (src[h.limbs - 1] >> 63) ? ~0ull : 0ull;
                for (int l = 0; l < limbs_cap; ++l) dst[l] = l < h.limbs ? src[l] : ext;
            }
    }
    return GI_OK;
}

}  // extern "C"
