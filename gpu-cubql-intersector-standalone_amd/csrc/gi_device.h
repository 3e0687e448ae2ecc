// gi_device.h -- launchers of the gfx950 kernels (implemented in
// gi_device.hip). All launches are asynchronous on the given stream.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gi {

constexpr uint32_t kSentinel = 0xFFFFFFFFu;
// Bit 31 of an internal node's left-child word: its right child is a leaf.
// With it the right child's index follows from the left's (Karras: children
// gamma and gamma + 1), so both can be loaded at once. Node indices stay
// below 2^31 (meshes of fewer than 2^30 triangles).
constexpr uint32_t kRightLeaf = 0x80000000u;
constexpr uint32_t kNodeIndexMask = 0x7FFFFFFFu;

// Morton precision of the BVH keys, bits per axis. Only the tree's quality
// depends on it (equal keys are ordered by triangle index); 8 bits per axis
// (24-bit keys, 3 sort passes) serve meshes up to 2^21 triangles as well as
// 10 bits (30-bit keys, 4 passes) and sort a pass faster.
inline int morton_axis_bits(uint32_t largest_mesh_triangles) {
    return largest_mesh_triangles <= (1u << 21) ? 8 : 10;
}

// Pipeline counters (device, uint32 each), cleared at the start of every
// intersection. Each sits alone in a 256-byte span: the appends of every
// stage are atomics on these words, and atomics on one line serialise, so
// counters sharing a line slowed each other's reservations (a filter launch
// spent 12 us on them). The enum values are word offsets; CNT_COUNT is the
// size of the block, and the two grid flags follow it.
constexpr uint32_t kCntStride = 64;
enum Counter : uint32_t {
    CNT_TRAVERSAL_HITS = 0 * kCntStride,    // FP32 box overlaps emitted by the traversal
    CNT_CANDIDATES = 1 * kCntStride,        // exact (double) box overlaps = candidate pairs
    CNT_FILTER_NO = 2 * kCntStride,
    CNT_FILTER_GREEN = 3 * kCntStride,
    CNT_FILTER_YELLOW = 4 * kCntStride,
    CNT_EXACT_LIST = 5 * kCntStride,        // pairs handed to the exact stage
    CNT_RESULTS = 6 * kCntStride,           // intersecting pairs written by the exact stage
    CNT_SEGMENTS = 7 * kCntStride,          // segments written by the exact stage
    CNT_EXACT_YELLOW_YES = 8 * kCntStride,  // YELLOW pairs the exact stage found intersecting
    CNT_EXACT_YELLOW_NO = 9 * kCntStride,
    CNT_COPLANAR = 10 * kCntStride,
    CNT_DEGENERATE = 11 * kCntStride,
    CNT_UNSOUND = 12 * kCntStride,          // filter said GREEN, exact said no (must stay 0)
    CNT_DEFERRED = 13 * kCntStride,         // coplanar pairs moved from the fast exact pass to the full one
    CNT_RESIDUAL = 14 * kCntStride,         // pairs left to the host residual stage (off-grid pairs that do not fit)
    CNT_FILTER_PRED = 15 * kCntStride,      // orient3d signs evaluated by the FP64 filter
    CNT_EXACT_PRED = 16 * kCntStride,       // orient3d and orient2d signs evaluated by the exact stage
    CNT_OVERFLOW = 17 * kCntStride,         // a count wrapped past 2^32 (the query is too large)
    CNT_SLOW = 18 * kCntStride,             // pairs the one-lane exact passes leave to the general ones
    CNT_COUNT = 19 * kCntStride
};

// Exact segments on the device: 8 numbers (x0 y0 z0 w0 x1 y1 z1 w1) of
// `limbs` 64-bit words each, two's complement: 2 limbs (int128) for grid
// meshes, 4 (256 bits) with a power-of-two scale per segment for general
// doubles. Pools are sized for the wider form.
constexpr int kSegLimbsGrid = 2;
constexpr int kSegLimbsWide = 4;
constexpr size_t kSegPoolWords = 8 * kSegLimbsWide;  // uint64 words per pool slot

// Rigid transform applied as CGAL's Aff_transformation_3 evaluates it:
// x' = ((m00*x + m01*y) + m02*z) + t0, every operation rounded separately.
struct Xform {
    double m[3][4];
    int identity;
    int snap;  // round the result to the nearest integer (grid meshes)
    double lo[3], hi[3];  // bounds of the source vertices: moved, their corners frame the Morton codes
};

struct MeshDev {
    uint32_t nv, nt;
    const double* src;   // source vertices (xyz), used when v64 is null (an unmoved mesh)
    const double4* v64;  // moved vertices used by the predicates, or null
    const int4* vi;      // their integer copy (valid when the mesh is on the grid); with v64 null, the source
                         // vertices' integer copy of an unmoved mesh on the grid, or null
    const uint32_t* tri; // 3 * nt
    const float4* nodes; // BVH, 2 float4 per node, 2*nt - 1 nodes (a walked mesh only)
    const uint32_t* order;  // triangle ids in Morton order (the sorted values)
    const float4* tbox;     // triangle boxes: lo, hi of triangle t at 2t, 2t + 1
};

void launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t s);

// What a mesh upload learns about its arrays on the device: the bounds of
// the vertices (NaN coordinates skipped; zeros for none), whether one of them
// is off the integer grid (not an integer within 2^26), and whether a
// triangle index is out of range.
struct MeshInspect {
    double lo[3], hi[3];
    uint32_t off_grid, bad_index;
};
// scratch: inspect_scratch_doubles() doubles; out: one MeshInspect
constexpr uint32_t kInspectBlocks = 512;
size_t inspect_scratch_doubles();
void launch_inspect_mesh(const double* xyz, uint32_t nv, const uint32_t* tri, uint32_t nt, double* scratch,
                         MeshInspect* out, hipStream_t s);
// A batch uploaded with per-mesh (local) indices: checks each against its
// mesh's vertex count (out->bad_index) and makes it global (adds the mesh's
// first vertex). Tiles as for launch_tri_boxes_forest.
void launch_forest_globalize(uint32_t* tri, const uint2* ttiles, uint32_t n_ttiles, const uint32_t* toff,
                             const uint32_t* voff, MeshInspect* out, hipStream_t s);
// grid flag = 1, bounds6 = empty box (order-preserving keys)
// zero counters[0, nzero) (one word per kCntStride) and the first
// kMsdZeroWords of each of msd_sorts bucket-sort word sets (msd_words +
// k kMsdWords), and set grid_flags[slot] = 1 for the slots in stage_mask
void launch_query_init(uint32_t* counters, uint32_t nzero, uint32_t* grid_flags, uint32_t stage_mask,
                       uint32_t* msd_words, uint32_t msd_sorts, hipStream_t s);
// Grid-stride width of the vertex part of a mesh's stage launch.
constexpr uint32_t kPrepareMaxBlocks = 4096;
// One mesh's staging: its vertices (moved by *xf, in device memory so that a
// captured launch sequence picks up a new transform from the copy before it;
// grid flag cleared when a vertex is off the grid) and its triangles (boxes,
// and with keys non-null the Morton keys, identity values and, with hist0
// non-null, the first radix pass's per-tile digit counts, as a SortJobU32
// takes them with hist0_ready). The Morton frame is the box of xf's moved
// source-bound corners (xf->lo, xf->hi).
struct StageJob {
    const double* src;
    uint32_t nv;
    const uint32_t* tri;
    uint32_t nt;
    const Xform* xf;
    double4* v64;
    int4* vi;
    uint32_t* grid_flag;
    int axis_bits;
    float4* box;
    uint32_t* keys;
    uint32_t* vals;
    uint32_t* hist0;
    // nv = 0 (no vertex blocks, no moved copies: an unmoved mesh) leaves the
    // grid flag as the query's init set it (1); force_off_grid clears it
    // (source vertices off the grid)
    int force_off_grid;
    // bucket sort (msd_sort_pair): the tiles add their counts of the top
    // kMsdBits key bits (key >> msd_shift) to msd_total (zeroed by the
    // query's init); null: none
    uint32_t* msd_total;
    int msd_shift;
    // an unmoved mesh on the grid: its source vertices as integers (one
    // 16-byte load per vertex instead of three 8-byte ones), or null
    const int4* vsrc;
};
// vsrc[i] = source vertex i as integers, for a mesh whose source vertices
// are all on the grid (integers within 2^26)
void launch_src_int4(const double* src, uint32_t nv, int4* vsrc, hipStream_t s);
// Up to two meshes staged in one launch.
void launch_stage_meshes(const StageJob* jobs, int njobs, hipStream_t s);

// Forest (a batch of meshes stored back to back, indices global): tiles of
// kForestTile vertices/triangles, each inside one mesh, tile = (mesh, first).
// mesh_vtile0[m]..mesh_vtile0[m+1] are mesh m's vertex tiles; partials needs
// 6 floats per vertex tile; bounds 6 keys per mesh. Vertices are used as
// given (no transform).
constexpr uint32_t kForestTile = 1024;
void launch_prepare_forest(const double* src, const uint2* vtiles, uint32_t n_vtiles, const uint32_t* voff,
                           const uint32_t* mesh_vtile0, uint32_t M, double4* v64, int4* vi, uint32_t* grid_flag,
                           float* partials, uint32_t* bounds, hipStream_t s);
// boxes and keys mesh << 3 axis_bits | Morton code (3 axis_bits + bits(M) <= 32;
// sort on that many bits)
void launch_tri_boxes_forest(const uint32_t* tri, const uint2* ttiles, uint32_t n_ttiles, const uint32_t* toff,
                             const double4* v64, const int4* vi, const uint32_t* grid_flag, const uint32_t* bounds,
                             int axis_bits, float4* box, uint32_t* keys, uint32_t* vals, hipStream_t s);

// LSD radix sort of (key, value) pairs, 8 bits per pass. tmp_* are ping-pong
// buffers of n elements; hist needs sort_hist_elems(n) entries, scan_tmp
// scan_tmp_elems(sort_hist_elems(n)). Returns 1 when the sorted data ends in
// the tmp buffers, 0 when it ends in (keys, vals).
size_t sort_hist_elems(size_t n);
size_t scan_tmp_elems(size_t n);
// whether a sort of n uint32 keys runs the multi-tile passes (and so can take
// the first pass's counts from launch_stage_mesh)
bool radix_sort_takes_hist0(uint32_t n);
// Two meshes' Morton sorts at once: each multi-tile pass of both runs in the
// same launches. in_tmp[j] = 1 when sort j's result ends in (tkeys, tvals).
struct SortJobU32 {
    uint32_t* keys; uint32_t* vals; uint32_t* tkeys; uint32_t* tvals;
    uint32_t n;
    int bits;
    bool hist0_ready;
    uint32_t* hist;      // sort_hist_elems(n)
    uint32_t* scan_tmp;  // scan_tmp_elems(sort_hist_elems(n))
};
void radix_sort_u32_pair(const SortJobU32* jobs, int njobs, hipStream_t s, int* in_tmp);
// Morton sorts of single meshes (13 to 24 key bits, more than one workgroup's
// keys) by buckets of the top kMsdBits bits, two launches for one or two
// sorts: launch_stage_mesh counts the buckets (StageJob.msd_total), then
// every key goes to its bucket (one reservation per bucket and tile) and
// each bucket is sorted by its low bits in LDS. Within a bucket equal low
// bits are equal keys, so no step needs a stable rank: the order among equal
// keys is whatever the atomics give (the tree over them is a valid Karras
// tree either way, and every result is ordered afterwards). The sorted keys
// end in (keys, vals).
//
// A bucket of more than kMsdBig keys (the first kMsdBigSlots of them per
// sort) is split: the scatter also counts its keys' low parts in a global
// histogram, and workgroups of kMsdSlice keys each reserve their places per
// low part with global atomics, instead of one workgroup walking the whole
// bucket (a UV sphere's pole cells hold ~15K keys).
constexpr int kMsdBits = 12;
constexpr uint32_t kMsdBuckets = 1u << kMsdBits;
constexpr uint32_t kMsdBig = 4096, kMsdSlice = 2048;
constexpr int kMsdBigSlots = 32;
// Words per sort: the totals and cursors (zeroed by the query's init), the
// starts and the buckets' slots (written by the scatter), the slot table
// (bucket of each slot, then the slots' first slice: 33 words), then the
// split buckets' done counters, low-part histograms and low-part cursors:
// zero when allocated, and every sort leaves them zero (the last slice of a
// bucket clears its slot).
constexpr uint32_t kMsdOffCursor = kMsdBuckets, kMsdOffStart = 2 * kMsdBuckets, kMsdOffSlot = 3 * kMsdBuckets,
                   kMsdOffTab = 4 * kMsdBuckets, kMsdOffDone = kMsdOffTab + 128, kMsdOffHist = kMsdOffDone + 128,
                   kMsdOffBigCursor = kMsdOffHist + kMsdBigSlots * kMsdBuckets,
                   kMsdWords = kMsdOffBigCursor + kMsdBigSlots * kMsdBuckets;
constexpr uint32_t kMsdZeroWords = 2 * kMsdBuckets;  // zeroed by every query's init
bool msd_sort_fits(uint32_t n, int bits);
struct MsdSortJob {
    uint32_t* keys; uint32_t* vals; uint32_t* tkeys; uint32_t* tvals;
    uint32_t n;
    int bits;
    uint32_t* words;  // kMsdWords
};
void msd_sort_pair(const MsdSortJob* jobs, int njobs, hipStream_t s);
int radix_sort_u64(uint64_t* keys, uint32_t* vals, uint64_t* tmp_keys, uint32_t* tmp_vals, uint32_t n, int bits,
                   uint32_t* hist, uint32_t* scan_tmp, hipStream_t s);
// exclusive scan of n uint32 (in place allowed); writes the total to *total
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* scan_tmp, uint32_t* total,
                        hipStream_t s);

// Karras LBVH over the sorted Morton keys, boxes included: leaf k (node
// n-1+k) holds prim sorted_prims[k] and its box; internal node i holds its
// box, its left child and its rope (next node in pre-order after its
// subtree); range[i] = its leaf range. `blocks` needs block_levels_elems(n)
// float4 (64-ary levels of leaf boxes, then a mask of the internal nodes
// whose range crosses a build chunk). Running it again on the same sorted keys with new prim boxes is the
// refit after a rigid motion (same topology, new boxes).
constexpr uint32_t kBuildChunk = 512;  // default sorted leaves per build workgroup (256 or 512)
constexpr int kMaxBlockLevels = 6;
struct BlockLevels {
    const float4* level[kMaxBlockLevels];
};
size_t block_levels_elems(uint32_t n);
// One tree to build (keys: Morton codes, or forest keys whose mesh ids lead,
// so that each mesh's BVH is a subtree of the forest).
template <class K>
struct BuildJob {
    const K* keys;
    const uint32_t* prims;
    uint32_t n;
    const float4* boxes;  // triangle boxes: lo, hi of triangle t at 2t, 2t + 1
    float4* nodes;
    uint2* range;
    float4* blocks;
};
// Builds one or two trees with one launch per kernel (the blocks of a launch
// are split between the trees); chunk_begin/chunk_end (may be null) bracket
// the k_build_chunks launch, recorded with event_flags (hipEventRecordExternal
// inside a graph capture, so that a replay times the launch).
void launch_build_bvh(const BuildJob<uint32_t>* jobs, int njobs, uint32_t chunk, hipStream_t s,
                      hipEvent_t chunk_begin, hipEvent_t chunk_end, unsigned event_flags);
// roots[m] = (subtree root of mesh m, rope leaving it)
void launch_forest_roots(const uint2* range, const float4* nodes, uint32_t n, const uint32_t* toff, uint32_t M,
                         uint2* roots, hipStream_t s);

// keysB/rootsA null: whole trees; else forests (sorted B keys whose mesh id
// starts at bit mesh_shift, A mesh roots)
// variant: launch shape of the traversal, 0..3 = (workgroup size, candidate
// pairs staged per wave) (64, 256), (128, 256), (256, 256), (256, 128)
void launch_traverse(MeshDev A, MeshDev B, const uint32_t* keysB, int mesh_shift, const uint2* rootsA, uint2* cand,
                     uint32_t cap, uint32_t* counters, int variant, hipStream_t s);
// blocks: workgroups of the grid-stride launch
void launch_filter(MeshDev A, MeshDev B, const uint2* cand, uint32_t cap, const uint32_t* grid_flags,
                   uint2* exact_list, uint32_t exact_cap, uint32_t* counters, int blocks, hipStream_t s);
// Exact stages a query launches: the grid flags are only known on the device,
// so the host launches both unless it knows which one they will select.
// kExactGridSlowLater: the passes for the pairs the one-lane passes leave
// (grid: 8-lane pass, then the deferred coplanar/degenerate pass; general
// doubles: the general one-lane pass and its deferred pass) are not queued
// here; the caller runs launch_exact_grid_slow / launch_exact_wide_slow when
// the counter read-back shows such pairs.
enum { kExactGrid = 1, kExactWide = 2, kExactGridSlowLater = 4 };
// defer: room for the coplanar pairs the fast pass hands to the full pass
// (at least as many entries as the exact list; the candidate queue, free by
// then, serves). Grid meshes run the int64/int128 stage; general doubles the
// wide stage (int128/256-bit), whose pairs that do not fit go to `resid`
// (resid_cap entries, the host residual stage). seg_pool holds kSegPoolWords
// words per slot, seg_exp one scale per slot (wide stage).
// slow: room for the pairs the one-lane passes leave (as many as exact_list)
void launch_exact(MeshDev A, MeshDev B, const uint2* exact_list, uint32_t exact_cap, const uint32_t* grid_flags,
                  int bitsB, uint64_t* rec_key, uint2* rec_seg, uint32_t rec_cap, uint64_t* seg_pool,
                  int32_t* seg_exp, uint32_t seg_cap, uint2* defer, uint32_t defer_cap, uint2* slow,
                  uint32_t slow_cap, uint2* resid, uint32_t resid_cap, uint32_t* counters, int blocks, int kinds,
                  hipStream_t s);
void launch_exact_wide_slow(MeshDev A, MeshDev B, const uint32_t* grid_flags, int bitsB, uint64_t* rec_key,
                            uint2* rec_seg, uint32_t rec_cap, uint64_t* seg_pool, int32_t* seg_exp, uint32_t seg_cap,
                            uint2* defer, uint32_t defer_cap, uint2* slow, uint32_t slow_cap, uint2* resid,
                            uint32_t resid_cap, uint32_t* counters, int blocks, hipStream_t s);
void launch_exact_grid_slow(MeshDev A, MeshDev B, const uint32_t* grid_flags, int bitsB, uint64_t* rec_key,
                            uint2* rec_seg, uint32_t rec_cap, uint64_t* seg_pool, uint32_t seg_cap, uint2* defer,
                            uint32_t defer_cap, uint2* slow, uint32_t slow_cap, uint32_t* counters, int blocks,
                            hipStream_t s);
void launch_iota(uint32_t* p, uint32_t n, hipStream_t s);
// Sorted output of up to output_small_max() result records in one launch,
// then their segments; the record count is read on the device, so this is
// queued before the counters reach the host. Does nothing unless
// output_small_fits(R, key_bits) and nothing overflowed.
uint32_t output_small_max();
bool output_small_fits(uint32_t R, int key_bits);
// Segments are gathered at the width the exact stage wrote (grid flags on
// the device: 2 limbs, else 4 with scales copied to seg_exp_out); seg_out
// holds out_slots segments of kSegPoolWords words. counters_out (host pinned
// memory as the device addresses it, or null) receives the counter words
// (counters[k * kCntStride]) and the two grid flags (counters[CNT_COUNT],
// [CNT_COUNT + 1]) at the same offsets, in place of a copy after the launches.
void launch_output_small(const uint64_t* rec_key, const uint2* rec_seg, const uint32_t* counters, uint32_t cap_rec,
                         uint32_t cap_seg, int key_bits, int bitsB, bool want_segments, uint2* pairs,
                         uint32_t* order, uint32_t* offs, uint64_t* seg_out, int32_t* seg_exp_out, double* seg_f64,
                         uint32_t out_slots, const uint64_t* seg_pool, const int32_t* seg_exp,
                         const uint32_t* grid_flags, uint32_t* counters_out, hipStream_t s);
void launch_gather_pairs(const uint64_t* sorted_key, const uint32_t* order, const uint2* rec_seg, uint32_t n,
                         int bitsB, uint2* pairs, uint32_t* nseg, hipStream_t s);
void launch_gather_segments(const uint32_t* order, const uint2* rec_seg, const uint32_t* seg_offsets, uint32_t n,
                            const uint64_t* seg_pool, const int32_t* seg_exp, int limbs, uint64_t* seg_out,
                            int32_t* seg_exp_out, double* seg_f64, hipStream_t s);

// forest results: pair_off[m] (M+1 entries) delimits mesh pair m's sorted
// pairs, whose indices become per-mesh
void launch_forest_localize(uint2* pairs, uint32_t R, const uint32_t* toffA, const uint32_t* toffB, uint32_t M,
                            uint32_t* pair_off, hipStream_t s);

}  // namespace gi
