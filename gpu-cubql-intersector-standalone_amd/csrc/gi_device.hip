// gi_device.hip -- gfx950 kernels of the mesh/mesh intersection path.
//
// Reference counterparts (files relative to the reference root):
//   prepare_vertices  src/common/utils.cu:479-593 (point intervals, rigid
//                     transform; here the transform is evaluated exactly as
//                     the CGAL path evaluates it, src/CPU/RotationTools.h:76-104)
//   tri_boxes         src/common/utils.cu:595-643 (FP64 -> FP32 boxes; here
//                     rounded outward instead of padded)
//   morton/sort/bvh   include/third-party/cubql/level_cut_builder.h:16-364
//                     and cuBQL's refit (replaced by a Karras LBVH with ropes)
//   traverse          src/single_tree/SingleTreeBatchCommon.cu:116-209 and
//                     include/third-party/cubql/fixedBoxQueryv2.h:55-105
//                     (count+fill with a 64-entry stack; here one stackless
//                     pass with wave-aggregated emission)
//   filter            src/GPU_exact_predicates/OriginShiftedFloatPredicate.cu:7-45
//   exact             src/GPU_exact_predicates/ExactBigIntPredicate.cu:381-548
//                     plus the host CGAL fallback and segment extraction
//   gather/sort       thrust::copy_if / sort in
//                     src/single_tree/SingleTreeBatchIntersector.cu:276-384
#include <hip/hip_runtime.h>

#include <string.h>

#include <type_traits>

#include "exact_tritri.h"
#include "exact_wide.h"
#include "filter.h"
#include "gi_device.h"
#include "gi_karras.h"

namespace gi {

#define GI_LAUNCH_CHECK() (void)hipGetLastError()

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// Lane (lane ^ M)'s copy of a float, every lane of the wave active: DPP
// quad permutes for M = 1, 2, LDS-free swizzles within 32 lanes for M = 4 ..
// 16, a permute through the LDS crossbar for 32.
template <int M>
__device__ __forceinline__ float lane_xor_f32(float x) {
    int v = __float_as_int(x);
    if constexpr (M == 1) v = __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);  // quad_perm [1, 0, 3, 2]
    else if constexpr (M == 2) v = __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);  // [2, 3, 0, 1]
    else if constexpr (M < 32) v = __builtin_amdgcn_ds_swizzle(v, (M << 10) | 0x1F);  // bit mode: lane ^ M
    else v = __shfl_xor(v, M);
    return __int_as_float(v);
}
// one butterfly step of a box (lo[3], hi[3]) at lane distance M: each lane
// then holds the union of its aligned block of 2M lanes
template <int M>
__device__ __forceinline__ void box_xor_step(float* b) {
#pragma unroll
    for (int c = 0; c < 3; ++c) b[c] = fminf(b[c], lane_xor_f32<M>(b[c]));
#pragma unroll
    for (int c = 3; c < 6; ++c) b[c] = fmaxf(b[c], lane_xor_f32<M>(b[c]));
}

// Runs of consecutive valid lanes of a wave with the same slot: one LDS
// atomic per run instead of one per lane (spatially coherent keys put long
// runs of one slot in a wave, and same-address LDS atomics serialise). A run
// head adds len to its slot's counter; every lane's place among the
// increments of its slot is then the head's old value (from lane `head`)
// plus (lane - head). The atomics of several items are issued before any
// result is read, so their latencies overlap.
struct WaveRun {
    uint32_t len;  // run length (run heads only; 0 elsewhere)
    int head;      // this lane's run head
};
__device__ __forceinline__ WaveRun wave_run(uint32_t slot, bool valid) {
    const uint32_t lane = lane_id();
    const uint32_t prev = (uint32_t)__shfl_up((int)slot, 1);
    const uint64_t vmask = __ballot(valid);
    const bool head = valid && (lane == 0 || !((vmask >> (lane - 1)) & 1ull) || prev != slot);
    const uint64_t heads = __ballot(head);
    const uint64_t upto = lane == 63 ? heads : heads & ((2ull << lane) - 1ull);  // heads at lanes <= lane
    WaveRun r;
    r.head = upto ? 63 - __clzll(upto) : (int)lane;
    r.len = 0;
    if (head) {  // the run ends at the next head or invalid lane
        const uint64_t after = lane == 63 ? 0ull : ((heads | ~vmask) & ~((2ull << lane) - 1ull));
        r.len = (after ? (uint32_t)__ffsll((long long)after) - 1u : 64u) - lane;
    }
    return r;
}
__device__ __forceinline__ uint32_t wave_run_place(uint32_t old, const WaveRun& r) {
    return (uint32_t)__shfl((int)old, r.head) + (lane_id() - (uint32_t)r.head);
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8), each XCD with its own L2. Renumbered so that XCD x takes one
// contiguous run of the G logical blocks, neighbouring work (adjacent chunks,
// neighbouring leaves along the Morton curve) shares one L2. A bijection on
// [0, G) for any G.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t G) {
    constexpr uint32_t X = 8;
    const uint32_t q = G / X, r = G % X, x = b % X;
    return x * q + min(x, r) + b / X;
}

// The same for a launch whose blocks [0, nA) serve one tree and the rest
// another: every XCD takes a contiguous run of each tree's blocks, its share
// of the first tree first (block b runs on XCD b % 8; of the blocks on XCD x,
// as many as [0, nA) holds for x go to the first tree). The two trees can
// differ in work per block and still load every XCD alike.
__device__ __forceinline__ uint32_t xcd_block_split(uint32_t b, uint32_t G, uint32_t nA, bool& second) {
    constexpr uint32_t X = 8;
    const uint32_t x = b % X, r = b / X;
    const uint32_t qA = nA > x ? (nA - 1 - x) / X + 1 : 0u;           // first-tree blocks on XCD x
    const uint32_t startA = x * (nA / X) + min(x, nA % X);            // their first block number
    second = r >= qA;
    if (!second) return startA + r;
    const uint32_t startG = x * (G / X) + min(x, G % X);              // XCD x's first block overall
    return startG - startA + (r - qA);                                // second tree's block number
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

static inline uint32_t blocks_for(size_t n, uint32_t threads) { return (uint32_t)((n + threads - 1) / threads); }

// ------------------------------------------------------------------ fill

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

void launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t s) {
    if (n == 0) return;
    uint32_t blocks = blocks_for(n, 256);
    if (blocks > 4096) blocks = 4096;
    k_fill_u32<<<blocks, 256, 0, s>>>(p, value, n);
    GI_LAUNCH_CHECK();
}

// Zero the counters in the first nzero words (one per kCntStride), and mark
// the grid flag of every slot in `stage_mask` (a slot about to be prepared)
// as "all on the grid" until one of its vertices proves otherwise.
__global__ void k_query_init(uint32_t* counters, uint32_t nzero, uint32_t* grid_flags, uint32_t stage_mask,
                             uint32_t* msd_words, uint32_t msd_sorts) {
    const uint32_t t = threadIdx.x;
    if (t * kCntStride < nzero) counters[t * kCntStride] = 0u;
    if (t < 2 && ((stage_mask >> t) & 1u)) grid_flags[t] = 1u;
    for (uint32_t k = 0; k < msd_sorts; ++k) {  // 16-byte stores (kMsdWords is a multiple of 4)
        uint4* w = (uint4*)(msd_words + (size_t)k * kMsdWords);
        for (uint32_t i = t; i < kMsdZeroWords / 4; i += blockDim.x) w[i] = make_uint4(0u, 0u, 0u, 0u);
    }
}
static_assert(kMsdWords % 4 == 0 && kMsdZeroWords % 4 == 0, "16-byte stores");

void launch_query_init(uint32_t* counters, uint32_t nzero, uint32_t* grid_flags, uint32_t stage_mask,
                       uint32_t* msd_words, uint32_t msd_sorts, hipStream_t s) {
    k_query_init<<<1, msd_sorts ? 1024 : 64, 0, s>>>(counters, nzero, grid_flags, stage_mask, msd_words, msd_sorts);
    GI_LAUNCH_CHECK();
}

// ------------------------------------------------------------ inspection
//
// One launch per uploaded mesh: blocks [0, kInspectBlocks) take the vertices
// (grid-stride; per-block bounds to `part`, the off-grid flag by one atomic
// per wave that saw such a vertex), the rest take the triangle indices (an
// out-of-range index raises bad_index the same way); then one block joins the
// partial bounds.

size_t inspect_scratch_doubles() { return 6 * (size_t)kInspectBlocks; }

__device__ __forceinline__ double wave_min_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ double wave_max_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

__global__ __launch_bounds__(256) void k_inspect_mesh(const double* __restrict__ xyz, uint32_t nv,
                                                      const uint32_t* __restrict__ tri, size_t nidx,
                                                      double* __restrict__ part, MeshInspect* out) {
    if (blockIdx.x < kInspectBlocks) {
        double l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
        bool off = false;
        for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < nv; i += kInspectBlocks * 256)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double x = xyz[3 * (size_t)i + k];
                l[k] = fmin(l[k], x);
                h[k] = fmax(h[k], x);
                off = off || !(x == rint(x) && fabs(x) <= 67108864.0);
            }
        if (__ballot(off) && lane_id() == 0) atomicOr(&out->off_grid, 1u);
        __shared__ double red[6][4];
        const uint32_t wave = threadIdx.x >> 6;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double mn = wave_min_d(l[k]), mx = wave_max_d(h[k]);
            if (lane_id() == 0) { red[k][wave] = mn; red[3 + k][wave] = mx; }
        }
        __syncthreads();
        if (threadIdx.x < 6) {
            const int k = threadIdx.x;
            double r = red[k][0];
            for (int w = 1; w < 4; ++w) r = k < 3 ? fmin(r, red[k][w]) : fmax(r, red[k][w]);
            part[6 * (size_t)blockIdx.x + k] = r;
        }
        return;
    }
    bool bad = false;
    const size_t nb = gridDim.x - kInspectBlocks;
    for (size_t i = (blockIdx.x - kInspectBlocks) * 256 + threadIdx.x; i < nidx; i += nb * 256) bad = bad || tri[i] >= nv;
    if (__ballot(bad) && lane_id() == 0) atomicOr(&out->bad_index, 1u);
}

__global__ __launch_bounds__(256) void k_inspect_join(const double* __restrict__ part, MeshInspect* out) {
    __shared__ double red[6][4];
    const uint32_t wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double mn = INFINITY, mx = -INFINITY;
        for (uint32_t b = threadIdx.x; b < kInspectBlocks; b += 256) {
            mn = fmin(mn, part[6 * (size_t)b + k]);
            mx = fmax(mx, part[6 * (size_t)b + 3 + k]);
        }
        mn = wave_min_d(mn);
        mx = wave_max_d(mx);
        if (lane_id() == 0) { red[k][wave] = mn; red[3 + k][wave] = mx; }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int k = threadIdx.x;
        double lo = fmin(fmin(red[k][0], red[k][1]), fmin(red[k][2], red[k][3]));
        double hi = fmax(fmax(red[3 + k][0], red[3 + k][1]), fmax(red[3 + k][2], red[3 + k][3]));
        const bool any = lo <= hi;  // no finite-or-infinite coordinate: zeros
        out->lo[k] = any ? lo : 0.0;
        out->hi[k] = any ? hi : 0.0;
    }
}

void launch_inspect_mesh(const double* xyz, uint32_t nv, const uint32_t* tri, uint32_t nt, double* scratch,
                         MeshInspect* out, hipStream_t s) {
    launch_fill_u32(&out->off_grid, 0u, 2, s);
    const size_t nidx = 3 * (size_t)nt;
    const uint32_t iblocks = nidx ? std::min<uint32_t>(blocks_for(nidx, 256), 2048u) : 0u;
    k_inspect_mesh<<<kInspectBlocks + iblocks, 256, 0, s>>>(xyz, nv, tri, nidx, scratch, out);
    k_inspect_join<<<1, 256, 0, s>>>(scratch, out);
    GI_LAUNCH_CHECK();
}

__global__ __launch_bounds__(256) void k_forest_globalize(uint32_t* __restrict__ tri, const uint2* __restrict__ tiles,
                                                          const uint32_t* __restrict__ toff,
                                                          const uint32_t* __restrict__ voff, MeshInspect* out) {
    const uint2 t = tiles[blockIdx.x];
    const uint32_t end = min(t.y + kForestTile, toff[t.x + 1]);
    const uint32_t v0 = voff[t.x], nvm = voff[t.x + 1] - v0;
    bool bad = false;
    for (uint32_t i = t.y + threadIdx.x; i < end; i += blockDim.x)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t idx = tri[3 * (size_t)i + k];
            bad = bad || idx >= nvm;
            tri[3 * (size_t)i + k] = idx + v0;
        }
    if (__ballot(bad) && lane_id() == 0) atomicOr(&out->bad_index, 1u);
}

void launch_forest_globalize(uint32_t* tri, const uint2* ttiles, uint32_t n_ttiles, const uint32_t* toff,
                             const uint32_t* voff, MeshInspect* out, hipStream_t s) {
    if (n_ttiles) k_forest_globalize<<<n_ttiles, 256, 0, s>>>(tri, ttiles, toff, voff, out);
    GI_LAUNCH_CHECK();
}

// ------------------------------------------------------- vertex preparation

__device__ __forceinline__ double xform_coord(const Xform& xf, int r, double x, double y, double z) {
    // CGAL Aff_transformation_3 on a Cartesian point, left to right, no FMA
    return __dadd_rn(__dadd_rn(__dadd_rn(__dmul_rn(xf.m[r][0], x), __dmul_rn(xf.m[r][1], y)), __dmul_rn(xf.m[r][2], z)),
                     xf.m[r][3]);
}

__device__ __forceinline__ float f32_down(double x) {
    float f = (float)x;
    if ((double)f > x) f = nextafterf(f, -INFINITY);
    return f;
}
__device__ __forceinline__ float f32_up(double x) {
    float f = (float)x;
    if ((double)f < x) f = nextafterf(f, INFINITY);
    return f;
}
// order-preserving map float -> uint32 (for atomic min/max)
__device__ __forceinline__ uint32_t f2key(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key2f(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// block-reduce a box; thread 0 writes it as 6 floats to out (one box per
// block: a second pass reduces them, no same-address atomics)
__device__ __forceinline__ void block_box_partial(const float* l, const float* h, float* out) {
    __shared__ float red[6][4];
    uint32_t wave = threadIdx.x >> 6;
    for (int k = 0; k < 3; ++k) {
        float mn = wave_min(l[k]), mx = wave_max(h[k]);
        if (lane_id() == 0) { red[k][wave] = mn; red[3 + k][wave] = mx; }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        int k = threadIdx.x;
        float r = red[k][0];
        for (int w = 1; w < 4; ++w) r = k < 3 ? fminf(r, red[k][w]) : fmaxf(r, red[k][w]);
        out[k] = r;
    }
}

// source coordinates as the query sees them: moved, and snapped for grid meshes
__device__ __forceinline__ void moved_coords(double sx, double sy, double sz, const Xform& xf, double& x, double& y,
                                             double& z) {
    x = sx;
    y = sy;
    z = sz;
    if (!xf.identity) {
        double X = xform_coord(xf, 0, x, y, z);
        double Y = xform_coord(xf, 1, x, y, z);
        double Z = xform_coord(xf, 2, x, y, z);
        x = X; y = Y; z = Z;
    }
    if (xf.snap) { x = rint(x); y = rint(y); z = rint(z); }
}

// vertex i as the query sees it: moved, and snapped for grid meshes
__device__ __forceinline__ void moved_vertex(const double* __restrict__ src, uint32_t i, const Xform& xf, double& x,
                                             double& y, double& z) {
    x = src[3 * (size_t)i];
    y = src[3 * (size_t)i + 1];
    z = src[3 * (size_t)i + 2];
    if (!xf.identity) {
        double X = xform_coord(xf, 0, x, y, z);
        double Y = xform_coord(xf, 1, x, y, z);
        double Z = xform_coord(xf, 2, x, y, z);
        x = X; y = Y; z = Z;
    }
    if (xf.snap) { x = rint(x); y = rint(y); z = rint(z); }
}

// One vertex: moved, the double and integer copies, the grid test and the
// running FP32 bounds (Morton frame only, so they need not be rounded
// outward).
__device__ __forceinline__ bool prepare_vertex(const double* __restrict__ src, uint32_t i, const Xform& xf,
                                               double4* __restrict__ v64, int4* __restrict__ vi, float* l, float* h) {
    double x, y, z;
    moved_vertex(src, i, xf, x, y, z);
    v64[i] = make_double4(x, y, z, 0.0);
    const double L = 67108864.0;  // 2^26
    bool g = x == rint(x) && y == rint(y) && z == rint(z) && fabs(x) <= L && fabs(y) <= L && fabs(z) <= L;
    vi[i] = g ? make_int4((int)x, (int)y, (int)z, 0) : make_int4(0, 0, 0, 0);
    float fx = (float)x, fy = (float)y, fz = (float)z;
    l[0] = fminf(l[0], fx); l[1] = fminf(l[1], fy); l[2] = fminf(l[2], fz);
    h[0] = fmaxf(h[0], fx); h[1] = fmaxf(h[1], fy); h[2] = fmaxf(h[2], fz);
    return g;
}

// Moves the vertices (when a transform is set), classifies the mesh as grid
// or not, writes the integer copy and per-block partial bounds (Morton frame).

// ---------------------------------------------------------- triangle boxes

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Morton code (3 x axis_bits bits, axis_bits <= 10) of a box centroid in the
// mesh bounds frame
__device__ __forceinline__ uint32_t morton_code(const float* bl, const float* bh, const float* bmin, const float* bext,
                                                int axis_bits) {
    const float cells = (float)(1u << axis_bits);
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        float c = 0.5f * bl[k] + 0.5f * bh[k];
        float u = (bext[k] > 0.f && isfinite(bext[k])) ? (c - bmin[k]) / bext[k] : 0.f;
        u = isfinite(u) ? u : 0.f;
        q[k] = (uint32_t)fminf(fmaxf(u * cells, 0.f), cells - 1.f);
    }
    return (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
}

// FP32 box of triangle i rounded outward from the exact vertex coordinates
// (so it contains the exact box). Grid meshes gather 16-byte integer
// vertices instead of 32-byte doubles.
__device__ __forceinline__ void tri_box(const uint32_t* __restrict__ tri, uint32_t i, bool grid,
                                        const double4* __restrict__ v, const int4* __restrict__ vi, float* bl,
                                        float* bh) {
    uint32_t i0 = tri[3 * (size_t)i], i1 = tri[3 * (size_t)i + 1], i2 = tri[3 * (size_t)i + 2];
    if (grid) {
        int4 a = vi[i0], b = vi[i1], c = vi[i2];
        bl[0] = f32_down((double)min(min(a.x, b.x), c.x));
        bl[1] = f32_down((double)min(min(a.y, b.y), c.y));
        bl[2] = f32_down((double)min(min(a.z, b.z), c.z));
        bh[0] = f32_up((double)max(max(a.x, b.x), c.x));
        bh[1] = f32_up((double)max(max(a.y, b.y), c.y));
        bh[2] = f32_up((double)max(max(a.z, b.z), c.z));
    } else {
        double4 a = v[i0], b = v[i1], c = v[i2];
        bl[0] = f32_down(fmin(fmin(a.x, b.x), c.x));
        bl[1] = f32_down(fmin(fmin(a.y, b.y), c.y));
        bl[2] = f32_down(fmin(fmin(a.z, b.z), c.z));
        bh[0] = f32_up(fmax(fmax(a.x, b.x), c.x));
        bh[1] = f32_up(fmax(fmax(a.y, b.y), c.y));
        bh[2] = f32_up(fmax(fmax(a.z, b.z), c.z));
    }
}

__device__ __forceinline__ void morton_frame(const uint32_t* __restrict__ bounds6, float* bmin, float* bext) {
    for (int k = 0; k < 3; ++k) {
        bmin[k] = key2f(bounds6[k]);
        bext[k] = key2f(bounds6[3 + k]) - bmin[k];
    }
}

// radix sort tiles: SORT_ITEMS keys per thread, 8-bit digits
constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 4;
constexpr uint32_t SORT_TILE = SORT_THREADS * SORT_ITEMS;
constexpr int RADIX = 256;

// Everything of up to two meshes before their Morton sorts, in one launch.
// Per mesh, vertex blocks prepare the vertices (move, snap, grid check,
// double and int copies; grid-stride) and triangle blocks take one sort tile
// of 1024 triangles each: box, Morton key, identity value and the tile's
// histogram of the first sort digit (the sort's first count pass). A
// triangle moves its three vertices itself, with the same arithmetic as the
// vertex blocks, so the two halves do not wait for each other; the Morton
// frame comes from the host (the box of the moved corners of the source
// vertices' bounds), not from a reduction in this launch; centroids outside
// it would be clamped to the border cells. The boxes use the doubles: for a
// grid mesh the integer vertices are the same values.
// Block ranges: [vertex blocks of mesh 0][vertex blocks of mesh 1][triangle
// tiles of mesh 0, padded to a multiple of 8][triangle tiles of mesh 1];
// every range starts at a multiple of 8 (block b runs on XCD b % 8), so the
// tiles of a mesh take XCD-contiguous tile numbers (xcd_block) and
// neighbouring tiles' first-pass count words share an L2. Fields are picked
// with selects on block-uniform flags.
struct StageJobs {
    StageJob j[2];
    uint32_t vb0, vb1, tb0, tb1;  // blocks of each range (tb0 padded)
};

__global__ __launch_bounds__(256) void k_stage_mesh(StageJobs J) {
    const uint32_t b = blockIdx.x;
    const bool vert = b < J.vb0 + J.vb1;
    const bool second = vert ? b >= J.vb0 : b >= J.vb0 + J.vb1 + J.tb0;
    const StageJob& q = second ? J.j[1] : J.j[0];
    const Xform xf = *q.xf;
    if (vert) {
        const uint32_t vb = second ? J.vb1 : J.vb0, b0 = second ? b - J.vb0 : b;
        bool grid = true;
        float l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = b0 * blockDim.x + threadIdx.x; i < q.nv; i += vb * blockDim.x)
            grid = prepare_vertex(q.src, i, xf, q.v64, q.vi, l, h) && grid;
        if (__ballot(!grid) && lane_id() == 0) atomicAnd(q.grid_flag, 0u);
        return;
    }
    const uint32_t t0 = J.vb0 + J.vb1 + (second ? J.tb0 : 0u);
    const uint32_t ntiles = (q.nt + SORT_TILE - 1) / SORT_TILE;
    const uint32_t span = second ? J.tb1 : J.tb0;  // padded range of this mesh
    const uint32_t tile = xcd_block(b - t0, span);
    if (tile >= ntiles) return;  // padding
    if (q.force_off_grid && tile == 0 && threadIdx.x == 0) *q.grid_flag = 0u;  // unmoved, source off the grid
    __shared__ uint32_t hs[4][RADIX];
    __shared__ uint32_t mb[kMsdBuckets];  // the tile's bucket counts (bucket sort)
    __shared__ float frame[6];  // Morton frame: lo[3], ext[3]
    if (q.hist0)
        for (int d = threadIdx.x; d < 4 * RADIX; d += blockDim.x) (&hs[0][0])[d] = 0u;
    if (q.msd_total)
        for (int d = threadIdx.x; d < (int)kMsdBuckets; d += blockDim.x) mb[d] = 0u;
    if (threadIdx.x == 0) {
        double lo[3] = {xf.lo[0], xf.lo[1], xf.lo[2]}, hi[3] = {xf.hi[0], xf.hi[1], xf.hi[2]};
        if (!xf.identity) {
            for (int k = 0; k < 3; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; }
            for (int c = 0; c < 8; ++c) {
                const double x = (c & 1) ? xf.hi[0] : xf.lo[0], y = (c & 2) ? xf.hi[1] : xf.lo[1],
                             z = (c & 4) ? xf.hi[2] : xf.lo[2];
                for (int k = 0; k < 3; ++k) {
                    const double v = xform_coord(xf, k, x, y, z);
                    lo[k] = fmin(lo[k], v);
                    hi[k] = fmax(hi[k], v);
                }
            }
        }
        for (int k = 0; k < 3; ++k) {
            frame[k] = (float)lo[k];
            frame[3 + k] = (float)hi[k] - (float)lo[k];
        }
    }
    __syncthreads();
    // all items' indices, then all their vertices, before any store: the
    // gathers of the tile's four triangles per thread are in flight together
    const uint32_t* __restrict__ tri = q.tri;
    const double* __restrict__ src = q.src;
    uint32_t id[SORT_ITEMS][3];
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it) {
        const uint32_t i = tile * SORT_TILE + it * 256 + threadIdx.x;
#pragma unroll
        for (int c = 0; c < 3; ++c) id[it][c] = i < q.nt ? tri[3 * (size_t)i + c] : 0u;
    }
    double raw[SORT_ITEMS][3][3];
    if (q.vsrc) {  // an unmoved mesh on the grid: one 16-byte load per vertex
#pragma unroll
        for (int it = 0; it < SORT_ITEMS; ++it)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int4 v = q.vsrc[id[it][c]];
                raw[it][c][0] = (double)v.x;
                raw[it][c][1] = (double)v.y;
                raw[it][c][2] = (double)v.z;
            }
    } else {
#pragma unroll
        for (int it = 0; it < SORT_ITEMS; ++it)
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int k = 0; k < 3; ++k) raw[it][c][k] = src[3 * (size_t)id[it][c] + k];
    }
    uint32_t bucket[SORT_ITEMS] = {};  // bucket sort: the items' buckets
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it) {
        const uint32_t i = tile * SORT_TILE + it * 256 + threadIdx.x;
        if (i >= q.nt) continue;
        double p[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c) moved_coords(raw[it][c][0], raw[it][c][1], raw[it][c][2], xf, p[c][0], p[c][1], p[c][2]);
        float bl[3], bh[3];
        for (int k = 0; k < 3; ++k) {
            bl[k] = f32_down(fmin(fmin(p[0][k], p[1][k]), p[2][k]));
            bh[k] = f32_up(fmax(fmax(p[0][k], p[1][k]), p[2][k]));
        }
        q.box[2 * (size_t)i] = make_float4(bl[0], bl[1], bl[2], 0.f);
        q.box[2 * (size_t)i + 1] = make_float4(bh[0], bh[1], bh[2], 0.f);
        if (q.keys) {
            const uint32_t key = morton_code(bl, bh, frame, frame + 3, q.axis_bits);
            q.keys[i] = key;
            q.vals[i] = i;
            if (q.hist0) atomicAdd(&hs[threadIdx.x >> 6][key & (RADIX - 1)], 1u);
            bucket[it] = key >> q.msd_shift;
        }
    }
    if (q.hist0) {
        __syncthreads();
        for (int d = threadIdx.x; d < RADIX; d += blockDim.x)
            q.hist0[(size_t)d * ntiles + tile] = hs[0][d] + hs[1][d] + hs[2][d] + hs[3][d];
    }
    if (q.msd_total) {  // the tile's bucket counts, then one global add per bucket it holds
#pragma unroll
        for (int it = 0; it < SORT_ITEMS; ++it) {
            const uint32_t i = tile * SORT_TILE + it * 256 + threadIdx.x;
            const WaveRun r = wave_run(bucket[it], i < q.nt);
            if (r.len) atomicAdd(&mb[bucket[it]], r.len);
        }
        __syncthreads();
        for (int d = threadIdx.x; d < (int)kMsdBuckets; d += blockDim.x)
            if (mb[d]) atomicAdd(&q.msd_total[d], mb[d]);
    }
}

__global__ __launch_bounds__(256) void k_src_int4(const double* __restrict__ src, uint32_t nv, int4* __restrict__ vsrc) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) {
        const double* p = src + 3 * (size_t)i;
        vsrc[i] = make_int4((int)p[0], (int)p[1], (int)p[2], 0);
    }
}

void launch_src_int4(const double* src, uint32_t nv, int4* vsrc, hipStream_t s) {
    if (nv == 0) return;
    k_src_int4<<<std::min<uint32_t>(blocks_for(nv, 256), 4096u), 256, 0, s>>>(src, nv, vsrc);
    GI_LAUNCH_CHECK();
}

void launch_stage_meshes(const StageJob* jobs, int njobs, hipStream_t s) {
    StageJobs J;
    memset(&J, 0, sizeof(J));
    uint32_t vb[2] = {0, 0}, tb[2] = {0, 0};
    for (int k = 0; k < njobs && k < 2; ++k) {
        J.j[k] = jobs[k];
        // multiples of 8 workgroups, so that every range starts on XCD 0
        vb[k] = jobs[k].nv ? (std::min<uint32_t>(blocks_for(jobs[k].nv, 256), kPrepareMaxBlocks) + 7u) & ~7u : 0u;
        tb[k] = (blocks_for(jobs[k].nt, SORT_TILE) + 7u) & ~7u;
    }
    if (njobs < 2) J.j[1] = J.j[0];
    J.vb0 = vb[0];
    J.vb1 = vb[1];
    J.tb0 = tb[0];
    J.tb1 = tb[1];
    const uint32_t blocks = vb[0] + vb[1] + tb[0] + tb[1];
    if (blocks == 0) return;
    k_stage_mesh<<<blocks, 256, 0, s>>>(J);
    GI_LAUNCH_CHECK();
}

__device__ __forceinline__ void box_absorb(const float4* __restrict__ e, size_t idx, float* b) {
    float4 l = e[2 * idx], h = e[2 * idx + 1];
    b[0] = fminf(b[0], l.x); b[1] = fminf(b[1], l.y); b[2] = fminf(b[2], l.z);
    b[3] = fmaxf(b[3], h.x); b[4] = fmaxf(b[4], h.y); b[5] = fmaxf(b[5], h.z);
}

// the union of the wave's boxes on every lane (every lane active)
__device__ __forceinline__ void wave_box_reduce(float* b) {
    box_xor_step<32>(b);
    box_xor_step<16>(b);
    box_xor_step<8>(b);
    box_xor_step<4>(b);
    box_xor_step<2>(b);
    box_xor_step<1>(b);
}

// ------------------------------------------------------ forests (batches)
//
// A batch stores its meshes back to back (mesh m's vertices at
// [voff[m], voff[m+1]), triangles at [toff[m], toff[m+1]), indices global).
// Work is cut into tiles of kForestTile elements that never straddle two
// meshes: tile t = (mesh, first element). Per-mesh results (bounds) come from
// per-tile partials, so no same-address atomics are needed.

__global__ __launch_bounds__(256) void k_prepare_vertices_forest(const double* __restrict__ src,
                                                                 const uint2* __restrict__ tiles,
                                                                 const uint32_t* __restrict__ voff,
                                                                 double4* __restrict__ v64, int4* __restrict__ vi,
                                                                 uint32_t* grid_flag, float* __restrict__ part) {
    const uint2 t = tiles[blockIdx.x];
    const uint32_t end = min(t.y + kForestTile, voff[t.x + 1]);
    Xform id;
    id.identity = 1;
    id.snap = 0;
    bool grid = true;
    float l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = t.y + threadIdx.x; i < end; i += blockDim.x) grid = prepare_vertex(src, i, id, v64, vi, l, h) && grid;
    uint64_t bad = __ballot(!grid);
    if (bad && lane_id() == 0) atomicAnd(grid_flag, 0u);
    block_box_partial(l, h, part + 6 * (size_t)blockIdx.x);
}

// one wave per mesh: union of its tiles' partial boxes -> 6 keys per mesh
__global__ __launch_bounds__(256) void k_forest_bounds(const float* __restrict__ part,
                                                       const uint32_t* __restrict__ mesh_tile0, uint32_t M,
                                                       uint32_t* __restrict__ bounds) {
    uint32_t m = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (m >= M) return;
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t t = mesh_tile0[m] + lane_id(); t < mesh_tile0[m + 1]; t += 64)
        for (int k = 0; k < 3; ++k) {
            b[k] = fminf(b[k], part[6 * (size_t)t + k]);
            b[3 + k] = fmaxf(b[3 + k], part[6 * (size_t)t + 3 + k]);
        }
    wave_box_reduce(b);
    if (lane_id() == 0)
        for (int k = 0; k < 6; ++k) bounds[6 * (size_t)m + k] = f2key(b[k]);
}

// Triangle boxes and keys (mesh id << 3 axis_bits | Morton code in the mesh's
// own frame, within 32 bits): one radix sort then groups the batch by mesh and
// orders each mesh along its Morton curve.
__global__ __launch_bounds__(256) void k_tri_boxes_forest(const uint32_t* __restrict__ tri,
                                                          const uint2* __restrict__ tiles,
                                                          const uint32_t* __restrict__ toff,
                                                          const double4* __restrict__ v, const int4* __restrict__ vi,
                                                          const uint32_t* __restrict__ grid_flag,
                                                          const uint32_t* __restrict__ bounds, int axis_bits,
                                                          float4* __restrict__ box,
                                                          uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint2 t = tiles[blockIdx.x];
    const uint32_t end = min(t.y + kForestTile, toff[t.x + 1]);
    const bool grid = *grid_flag != 0;
    float bmin[3], bext[3];
    morton_frame(bounds + 6 * (size_t)t.x, bmin, bext);
    for (uint32_t i = t.y + threadIdx.x; i < end; i += blockDim.x) {
        float bl[3], bh[3];
        tri_box(tri, i, grid, v, vi, bl, bh);
        box[2 * (size_t)i] = make_float4(bl[0], bl[1], bl[2], 0.f);
        box[2 * (size_t)i + 1] = make_float4(bh[0], bh[1], bh[2], 0.f);
        keys[i] = (t.x << (3 * axis_bits)) | morton_code(bl, bh, bmin, bext, axis_bits);
        vals[i] = i;
    }
}

void launch_prepare_forest(const double* src, const uint2* vtiles, uint32_t n_vtiles, const uint32_t* voff,
                           const uint32_t* mesh_vtile0, uint32_t M, double4* v64, int4* vi, uint32_t* grid_flag,
                           float* partials, uint32_t* bounds, hipStream_t s) {
    launch_fill_u32(grid_flag, 1u, 1, s);
    if (n_vtiles == 0 || M == 0) return;
    k_prepare_vertices_forest<<<n_vtiles, 256, 0, s>>>(src, vtiles, voff, v64, vi, grid_flag, partials);
    k_forest_bounds<<<blocks_for(M, 4), 256, 0, s>>>(partials, mesh_vtile0, M, bounds);
    GI_LAUNCH_CHECK();
}

void launch_tri_boxes_forest(const uint32_t* tri, const uint2* ttiles, uint32_t n_ttiles, const uint32_t* toff,
                             const double4* v64, const int4* vi, const uint32_t* grid_flag, const uint32_t* bounds,
                             int axis_bits, float4* box, uint32_t* keys, uint32_t* vals, hipStream_t s) {
    if (n_ttiles == 0) return;
    k_tri_boxes_forest<<<n_ttiles, 256, 0, s>>>(tri, ttiles, toff, v64, vi, grid_flag, bounds, axis_bits, box, keys,
                                                vals);
    GI_LAUNCH_CHECK();
}

// -------------------------------------------------------------------- scan

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr uint32_t SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// block sums of a scan, and at least the 256 digit totals of a sort pass
size_t scan_tmp_elems(size_t n) {
    size_t b = (n + SCAN_TILE - 1) / SCAN_TILE + 1;
    return b < 257 ? 257 : b;
}

__device__ __forceinline__ uint32_t block_exclusive_sum(uint32_t v, uint32_t* total) {
    __shared__ uint32_t ws[SCAN_THREADS / 64];
    uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) ws[wave] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (uint32_t w = 0; w < SCAN_THREADS / 64; ++w) {
        if (w < wave) off += ws[w];
        tot += ws[w];
    }
    __syncthreads();
    if (total) *total = tot;
    return off + x - v;
}

// Two exclusive sums over the block in one pass (one pair of barriers):
// returns a's, sets b's exclusive sum and b's block total.
__device__ __forceinline__ uint32_t block_exclusive_sum2(uint32_t a, uint32_t b, uint32_t& b_excl, uint32_t& b_total) {
    __shared__ uint2 ws2[SCAN_THREADS / 64];
    uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    uint32_t x = a, y = b;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t xa = __shfl_up(x, o), yb = __shfl_up(y, o);
        if (lane >= (uint32_t)o) {
            x += xa;
            y += yb;
        }
    }
    if (lane == 63) ws2[wave] = make_uint2(x, y);
    __syncthreads();
    uint32_t offa = 0, offb = 0, totb = 0;
    for (uint32_t w = 0; w < SCAN_THREADS / 64; ++w) {
        const uint2 q = ws2[w];
        if (w < wave) {
            offa += q.x;
            offb += q.y;
        }
        totb += q.y;
    }
    __syncthreads();
    b_excl = offb + y - b;
    b_total = totb;
    return offa + x - a;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const uint32_t* __restrict__ in, uint32_t n,
                                                              uint32_t* __restrict__ sums) {
    uint32_t base = blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        uint32_t i = base + it * SCAN_THREADS + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t tot;
    block_exclusive_sum(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_top(uint32_t* sums, uint32_t nb, uint32_t* total) {
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += SCAN_TILE) {
        uint32_t v[SCAN_ITEMS];
        uint32_t s = 0;
        for (int it = 0; it < SCAN_ITEMS; ++it) {
            uint32_t i = base + threadIdx.x * SCAN_ITEMS + it;
            v[it] = i < nb ? sums[i] : 0;
            s += v[it];
        }
        uint32_t tot;
        uint32_t off = block_exclusive_sum(s, &tot) + carry;
        for (int it = 0; it < SCAN_ITEMS; ++it) {
            uint32_t i = base + threadIdx.x * SCAN_ITEMS + it;
            if (i < nb) sums[i] = off;
            off += v[it];
        }
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                            uint32_t n, const uint32_t* __restrict__ sums) {
    uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        uint32_t i = base + it;
        v[it] = i < n ? in[i] : 0;
        s += v[it];
    }
    uint32_t off = block_exclusive_sum(s, nullptr) + sums[blockIdx.x];
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        uint32_t i = base + it;
        if (i < n) out[i] = off;
        off += v[it];
    }
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* scan_tmp, uint32_t* total,
                        hipStream_t s) {
    uint32_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (n == 0) {
        if (total) launch_fill_u32(total, 0, 1, s);
        return;
    }
    k_scan_reduce<<<nb, SCAN_THREADS, 0, s>>>(in, n, scan_tmp);
    k_scan_top<<<1, SCAN_THREADS, 0, s>>>(scan_tmp, nb, total);
    k_scan_down<<<nb, SCAN_THREADS, 0, s>>>(in, out, n, scan_tmp);
    GI_LAUNCH_CHECK();
}

// -------------------------------------------------------------- radix sort

size_t sort_hist_elems(size_t n) { return RADIX * ((n + SORT_TILE - 1) / SORT_TILE) + 1; }

// One pass of up to two independent sorts (two meshes' Morton keys) in each
// launch: upsweep and downsweep blocks [0, split) serve sort a, the rest sort
// b; digit-scan blocks [0, scan_split) sort a. Fields are picked with selects
// on a block-uniform flag (no dynamic indexing of the kernel arguments).
template <class K>
struct SortPass {
    const K* kin_a; const K* kin_b;
    const uint32_t* vin_a; const uint32_t* vin_b;
    K* kout_a; K* kout_b;
    uint32_t* vout_a; uint32_t* vout_b;
    uint32_t* hist_a; uint32_t* hist_b;      // per-tile digit counts, digit-major (d * ntiles + tile)
    uint32_t* totals_a; uint32_t* totals_b;  // digit totals
    uint32_t n_a, n_b, ntiles_a, ntiles_b;
    uint32_t split, scan_split;
    int shift;
};

// The tile loops take XCD-contiguous tile numbers (xcd_block): the count words
// of neighbouring tiles (hist is digit-major) and their runs of one digit in
// the sorted output then fill whole lines in one L2 instead of partial lines
// in eight (16.4 -> 2.0 MB written per upsweep of 2^21 keys, 31 -> 16.5 MB per
// downsweep).
// Lanes of a wave holding the same 8-bit digit (among the valid ones).
__device__ __forceinline__ uint64_t digit_match(uint32_t d, bool valid) {
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const uint64_t bb = __ballot((d >> b) & 1u);
        match &= ((d >> b) & 1u) ? bb : ~bb;
    }
    return match;
}

// Per-tile digit counts. A wave counts its keys item by item: the lowest lane
// of each group of equal digits adds the group's size to the wave's row, so
// no two lanes update one LDS word at once (LDS atomics on Morton digits,
// many lanes to one digit, serialised).
template <class K>
__global__ __launch_bounds__(SORT_THREADS) void k_sort_upsweep(SortPass<K> P) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const bool b = blk >= P.split;
    const K* __restrict__ keys = b ? P.kin_b : P.kin_a;
    uint32_t* __restrict__ hist = b ? P.hist_b : P.hist_a;
    const uint32_t n = b ? P.n_b : P.n_a, ntiles = b ? P.ntiles_b : P.ntiles_a;
    const uint32_t tile = blk - (b ? P.split : 0u);
    const int shift = P.shift;
    __shared__ uint32_t h[SORT_THREADS / 64][RADIX];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t base = tile * SORT_TILE + wave * (64 * SORT_ITEMS) + lane_id();
    K k[SORT_ITEMS];
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it) {
        const uint32_t i = base + it * 64;
        k[it] = i < n ? keys[i] : (K)0;
    }
    for (int d = threadIdx.x; d < (SORT_THREADS / 64) * RADIX; d += SORT_THREADS) (&h[0][0])[d] = 0;
    __syncthreads();
    const uint64_t lt = lanemask_lt();
    for (int it = 0; it < SORT_ITEMS; ++it) {
        const bool valid = base + it * 64 < n;
        const uint32_t d = (uint32_t)(k[it] >> shift) & (RADIX - 1);
        const uint64_t m = digit_match(d, valid);
        if (valid && (m & lt) == 0) h[wave][d] += (uint32_t)__popcll(m);
        wave_sync_lds();
    }
    __syncthreads();
    for (int d = threadIdx.x; d < RADIX; d += SORT_THREADS) {
        uint32_t c = 0;
        for (int w = 0; w < SORT_THREADS / 64; ++w) c += h[w][d];
        hist[(size_t)d * ntiles + tile] = c;
    }
}

// One block per digit: exclusive scan of that digit's per-tile counts (in
// place) and the digit total. The digit totals are scanned by each
// downsweep block, so a pass is three launches.
template <class K>
__global__ __launch_bounds__(SCAN_THREADS) void k_sort_digit_scan(SortPass<K> P) {
    const bool b = blockIdx.x >= P.scan_split;
    const uint32_t ntiles = b ? P.ntiles_b : P.ntiles_a;
    uint32_t* totals = b ? P.totals_b : P.totals_a;
    const uint32_t digit = blockIdx.x - (b ? P.scan_split : 0u);
    uint32_t* h = (b ? P.hist_b : P.hist_a) + (size_t)digit * ntiles;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += SCAN_TILE) {
        uint32_t v[SCAN_ITEMS];
        uint32_t s = 0;
        for (int it = 0; it < SCAN_ITEMS; ++it) {
            uint32_t i = base + threadIdx.x * SCAN_ITEMS + it;
            v[it] = i < ntiles ? h[i] : 0;
            s += v[it];
        }
        uint32_t tot;
        uint32_t off = block_exclusive_sum(s, &tot) + carry;
        for (int it = 0; it < SCAN_ITEMS; ++it) {
            uint32_t i = base + threadIdx.x * SCAN_ITEMS + it;
            if (i < ntiles) h[i] = off;
            off += v[it];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) totals[digit] = carry;
}

// A tile's keys and values into registers: wave w holds the contiguous
// segment [w * 64 * SORT_ITEMS, (w + 1) * 64 * SORT_ITEMS) of the tile, item
// by item.
template <class K>
__device__ __forceinline__ void tile_load(const K* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t n,
                                          uint32_t tile, K (&k)[SORT_ITEMS], uint32_t (&v)[SORT_ITEMS]) {
    const uint32_t base = tile * SORT_TILE + (threadIdx.x >> 6) * (64 * SORT_ITEMS) + lane_id();
#pragma unroll
    for (int it = 0; it < SORT_ITEMS; ++it) {
        const uint32_t i = base + it * 64;
        k[it] = i < n ? kin[i] : (K)0;
        v[it] = i < n ? vin[i] : 0u;
    }
}

// Stable rank + scatter of one tile (k, v as tile_load reads them): keys are
// ranked in (wave, item, lane) order, their original order. The tile is first
// put in digit order in LDS (key of digit d at the tile's start of d + the
// count of d in earlier waves + its rank in its wave), then written out in
// that order, so consecutive lanes write consecutive addresses of one digit's
// run (scattered from registers, a store of 64 lanes touched up to 64 lines).
// wc must be zero on entry; gbase must be complete (a barrier follows the
// first use).
template <class K>
__device__ __forceinline__ void tile_rank_scatter(const K (&k)[SORT_ITEMS], const uint32_t (&v)[SORT_ITEMS],
                                                  K* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n,
                                                  int shift, uint32_t tile, uint32_t (*wc)[RADIX],
                                                  const uint32_t* gbase, uint32_t* lstart, K* lk, uint32_t* lv) {
    static_assert(RADIX == SORT_THREADS, "one digit per thread");
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const uint32_t base = tile * SORT_TILE + wave * (64 * SORT_ITEMS);
    const uint64_t lt = lanemask_lt();
    uint32_t rank[SORT_ITEMS];
    for (int it = 0; it < SORT_ITEMS; ++it) {
        const bool valid = base + it * 64 + lane < n;
        const uint32_t d = (uint32_t)(k[it] >> shift) & (RADIX - 1);
        const uint64_t match = digit_match(d, valid);
        const uint32_t peers = (uint32_t)__popcll(match & lt);
        const uint32_t cur = wc[wave][d];
        rank[it] = cur + peers;
        wave_sync_lds();
        if (valid && peers == 0) wc[wave][d] = cur + (uint32_t)__popcll(match);
        wave_sync_lds();
    }
    __syncthreads();
    {
        // digit d (this thread): exclusive prefix over the waves, and the
        // tile's start of d (exclusive prefix of the tile's digit counts)
        const uint32_t d = threadIdx.x;
        uint32_t run = 0;
        for (int w = 0; w < SORT_THREADS / 64; ++w) {
            const uint32_t t = wc[w][d];
            wc[w][d] = run;
            run += t;
        }
        lstart[d] = block_exclusive_sum(run, nullptr);
    }
    __syncthreads();
    for (int it = 0; it < SORT_ITEMS; ++it) {
        if (base + it * 64 + lane >= n) continue;
        const uint32_t d = (uint32_t)(k[it] >> shift) & (RADIX - 1);
        const uint32_t lp = lstart[d] + wc[wave][d] + rank[it];
        lk[lp] = k[it];
        lv[lp] = v[it];
    }
    __syncthreads();
    const uint32_t cnt = min(SORT_TILE, n - tile * SORT_TILE);
    for (uint32_t j = threadIdx.x; j < cnt; j += SORT_THREADS) {
        const K key = lk[j];
        const uint32_t d = (uint32_t)(key >> shift) & (RADIX - 1);
        const uint32_t pos = gbase[d] + (j - lstart[d]);
        kout[pos] = key;
        vout[pos] = lv[j];
    }
}

template <class K>
__global__ __launch_bounds__(SORT_THREADS) void k_sort_downsweep(SortPass<K> P) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const bool b = blk >= P.split;
    const K* __restrict__ kin = b ? P.kin_b : P.kin_a;
    const uint32_t* __restrict__ vin = b ? P.vin_b : P.vin_a;
    K* __restrict__ kout = b ? P.kout_b : P.kout_a;
    uint32_t* __restrict__ vout = b ? P.vout_b : P.vout_a;
    const uint32_t* __restrict__ hist = b ? P.hist_b : P.hist_a;
    const uint32_t* __restrict__ totals = b ? P.totals_b : P.totals_a;
    const uint32_t n = b ? P.n_b : P.n_a, ntiles = b ? P.ntiles_b : P.ntiles_a;
    const uint32_t tile = blk - (b ? P.split : 0u);
    __shared__ uint32_t wc[SORT_THREADS / 64][RADIX];
    __shared__ uint32_t gbase[RADIX];
    __shared__ uint32_t lstart[RADIX];
    __shared__ K lk[SORT_TILE];
    __shared__ uint32_t lv[SORT_TILE];
    // the tile's keys first: their loads overlap those of the digit counts
    K k[SORT_ITEMS];
    uint32_t v[SORT_ITEMS];
    tile_load<K>(kin, vin, n, tile, k, v);
    for (int d = threadIdx.x; d < (SORT_THREADS / 64) * RADIX; d += SORT_THREADS) (&wc[0][0])[d] = 0;
    {
        // digit base = exclusive prefix of the digit totals (RADIX == SORT_THREADS)
        uint32_t t = totals[threadIdx.x];
        uint32_t ex = block_exclusive_sum(t, nullptr);
        gbase[threadIdx.x] = ex + hist[(size_t)threadIdx.x * ntiles + tile];
    }
    __syncthreads();
    tile_rank_scatter<K>(k, v, kout, vout, n, P.shift, tile, wc, gbase, lstart, lk, lv);
}

// Whole LSD radix sort in one workgroup (one launch) for inputs up to
// small_sort_max keys, where multi-block passes are bound by launch gaps. Every
// thread keeps SMALL_ITEMS keys in registers; wave w owns the contiguous
// segment [w * 64 * SMALL_ITEMS, (w + 1) * 64 * SMALL_ITEMS) and ranks it item
// by item (ballot match per 64 keys, running per-digit counts in LDS), so the
// order (wave, item, lane) is the input order and every pass is stable. A pass
// scatters through global memory (ping-pong buffers) and reloads after a
// workgroup barrier. Values wait in LDS, keys and ranks in registers.
constexpr int SMALL_THREADS = 1024;
// Up to 4 keys per thread: beyond 4,096 keys one workgroup (one CU) takes
// longer per pass than the three launches of a multi-tile pass.
template <class K>
constexpr int small_items() { return 4; }
template <class K>
constexpr uint32_t small_sort_max() { return SMALL_THREADS * small_items<K>(); }

template <class K>
__global__ __launch_bounds__(SMALL_THREADS) void k_sort_small(K* k0, uint32_t* v0, K* k1, uint32_t* v1, uint32_t n,
                                                              int passes) {
    constexpr int SMALL_ITEMS = small_items<K>();
    constexpr uint32_t kSmallSortMax = small_sort_max<K>();
    __shared__ uint32_t base[RADIX];
    __shared__ uint32_t wc[SMALL_THREADS / 64][RADIX];
    __shared__ uint32_t sval[kSmallSortMax];
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const uint64_t lt = lanemask_lt();
    const uint32_t seg = wave * 64 * SMALL_ITEMS + lane;
    K key[SMALL_ITEMS];
    uint32_t rank2[SMALL_ITEMS / 2];  // ranks < 2^16, two per register
    for (int it = 0; it < SMALL_ITEMS; ++it) {
        uint32_t i = seg + 64 * it;
        key[it] = i < n ? k0[i] : (K)0;
        if (i < n) sval[i] = v0[i];
    }
    for (int p = 0; p < passes; ++p) {
        K* kd = (p & 1) ? k0 : k1;
        uint32_t* vd = (p & 1) ? v0 : v1;
        const int shift = 8 * p;
        for (int d = threadIdx.x; d < (SMALL_THREADS / 64) * RADIX; d += SMALL_THREADS) (&wc[0][0])[d] = 0;
        for (int q = 0; q < SMALL_ITEMS / 2; ++q) rank2[q] = 0;
        __syncthreads();
        for (int it = 0; it < SMALL_ITEMS; ++it) {
            bool valid = seg + 64 * it < n;
            uint32_t d = (uint32_t)(key[it] >> shift) & (RADIX - 1);
            uint64_t match = __ballot(valid);
            for (int b = 0; b < 8; ++b) {
                uint64_t bb = __ballot((d >> b) & 1u);
                match &= ((d >> b) & 1u) ? bb : ~bb;
            }
            uint32_t peers = (uint32_t)__popcll(match & lt);
            uint32_t cur = wc[wave][d];
            rank2[it >> 1] |= (cur + peers) << (16 * (it & 1));
            wave_sync_lds();
            if (valid && peers == 0) wc[wave][d] = cur + (uint32_t)__popcll(match);
            wave_sync_lds();
        }
        __syncthreads();
        if (threadIdx.x < RADIX) {  // per digit: exclusive prefix over the waves, total to base
            uint32_t run = 0;
            for (int w = 0; w < SMALL_THREADS / 64; ++w) {
                uint32_t t = wc[w][threadIdx.x];
                wc[w][threadIdx.x] = run;
                run += t;
            }
            base[threadIdx.x] = run;
        }
        __syncthreads();
        if (threadIdx.x < 64) {  // exclusive scan of the 256 digit totals by one wave
            uint32_t c[4], s = 0;
            for (int q = 0; q < 4; ++q) { c[q] = base[4 * threadIdx.x + q]; s += c[q]; }
            uint32_t x = s;
            for (int o = 1; o < 64; o <<= 1) {
                uint32_t y = __shfl_up(x, o);
                if (lane >= (uint32_t)o) x += y;
            }
            uint32_t off = x - s;
            for (int q = 0; q < 4; ++q) { base[4 * threadIdx.x + q] = off; off += c[q]; }
        }
        __syncthreads();
        for (int it = 0; it < SMALL_ITEMS; ++it) {
            if (seg + 64 * it >= n) continue;
            uint32_t d = (uint32_t)(key[it] >> shift) & (RADIX - 1);
            uint32_t pos = base[d] + wc[wave][d] + ((rank2[it >> 1] >> (16 * (it & 1))) & 0xFFFFu);
            kd[pos] = key[it];
            vd[pos] = sval[seg + 64 * it];
        }
        __syncthreads();
        if (p + 1 < passes)
            for (int it = 0; it < SMALL_ITEMS; ++it) {
                uint32_t i = seg + 64 * it;
                key[it] = i < n ? kd[i] : (K)0;
                if (i < n) sval[i] = vd[i];
            }
    }
}

// One sort's state across passes: ping-pong between (keys, vals) and (tk, tv).
template <class K>
struct SortRun {
    K* keys; uint32_t* vals; K* tk; uint32_t* tv;
    uint32_t n, ntiles;
    int passes;
    bool hist0_ready;  // the first pass's per-tile counts are already in hist
    uint32_t* hist; uint32_t* totals;
};

// The multi-tile passes of one or two sorts; pass p of both runs in the same
// three launches (a sort with fewer passes drops out). Returns, per sort, 1
// when the sorted data ends in (tk, tv).
template <class K>
static void multi_tile_passes(SortRun<K>* r, int nr, hipStream_t s, int* parity) {
    int maxp = 0;
    for (int j = 0; j < nr; ++j) maxp = std::max(maxp, r[j].passes);
    for (int p = 0; p < maxp; ++p) {
        SortPass<K> P{};
        P.shift = 8 * p;
        bool act[2] = {false, false};
        for (int j = 0; j < nr; ++j) act[j] = p < r[j].passes;
        auto fill = [&](int j, bool second) {
            const SortRun<K>& q = r[j];
            const bool odd = p & 1;
            const K* ki = odd ? q.tk : q.keys;
            const uint32_t* vi = odd ? q.tv : q.vals;
            K* ko = odd ? q.keys : q.tk;
            uint32_t* vo = odd ? q.vals : q.tv;
            if (!second) {
                P.kin_a = ki; P.vin_a = vi; P.kout_a = ko; P.vout_a = vo;
                P.hist_a = q.hist; P.totals_a = q.totals; P.n_a = q.n; P.ntiles_a = q.ntiles;
            } else {
                P.kin_b = ki; P.vin_b = vi; P.kout_b = ko; P.vout_b = vo;
                P.hist_b = q.hist; P.totals_b = q.totals; P.n_b = q.n; P.ntiles_b = q.ntiles;
            }
        };
        // slot a: the first active sort, slot b: the second (if any)
        int ja = -1, jb = -1;
        for (int j = 0; j < nr; ++j)
            if (act[j]) { if (ja < 0) ja = j; else jb = j; }
        if (ja < 0) break;
        fill(ja, false);
        if (jb >= 0) fill(jb, true);
        // upsweep: only the sorts whose counts for this pass are not there yet
        {
            SortPass<K> U{};
            U.shift = P.shift;
            uint32_t nblk[2] = {0u, 0u};
            int k = 0;
            for (int j : {ja, jb}) {
                if (j < 0 || (p == 0 && r[j].hist0_ready)) continue;
                const K* ki = (p & 1) ? r[j].tk : r[j].keys;
                if (k == 0) { U.kin_a = ki; U.hist_a = r[j].hist; U.n_a = r[j].n; U.ntiles_a = r[j].ntiles; }
                else { U.kin_b = ki; U.hist_b = r[j].hist; U.n_b = r[j].n; U.ntiles_b = r[j].ntiles; }
                nblk[k++] = r[j].ntiles;
            }
            U.split = nblk[0];
            if (nblk[0] + nblk[1]) k_sort_upsweep<K><<<nblk[0] + nblk[1], SORT_THREADS, 0, s>>>(U);
        }
        P.split = P.ntiles_a;
        P.scan_split = RADIX;
        const uint32_t scans = jb >= 0 ? 2 * RADIX : RADIX;
        k_sort_digit_scan<K><<<scans, SCAN_THREADS, 0, s>>>(P);
        k_sort_downsweep<K><<<P.ntiles_a + (jb >= 0 ? P.ntiles_b : 0u), SORT_THREADS, 0, s>>>(P);
    }
    for (int j = 0; j < nr; ++j) parity[j] = r[j].passes & 1;
    GI_LAUNCH_CHECK();
}

// Ping-pongs between (keys, vals) and (tk, tv); returns 1 when the sorted
// data ends in (tk, tv), 0 when it ends in (keys, vals).
template <class K>
static int radix_sort_impl(K* keys, uint32_t* vals, K* tk, uint32_t* tv, uint32_t n, int bits, uint32_t* hist,
                           uint32_t* scan_tmp, hipStream_t s, bool hist0_ready = false) {
    if (n <= 1) return 0;
    if (n <= small_sort_max<K>()) {
        int passes = (bits + 7) / 8;
        k_sort_small<K><<<1, SMALL_THREADS, 0, s>>>(keys, vals, tk, tv, n, passes);
        GI_LAUNCH_CHECK();
        return passes & 1;
    }
    SortRun<K> r{keys, vals, tk, tv, n, (n + SORT_TILE - 1) / SORT_TILE, (bits + 7) / 8, hist0_ready, hist, scan_tmp};
    int parity = 0;
    multi_tile_passes<K>(&r, 1, s, &parity);
    return parity;
}


bool radix_sort_takes_hist0(uint32_t n) { return n > small_sort_max<uint32_t>(); }

void radix_sort_u32_pair(const SortJobU32* jobs, int njobs, hipStream_t s, int* in_tmp) {
    SortRun<uint32_t> r[2];
    int map[2], nr = 0;
    for (int j = 0; j < njobs; ++j) {
        const SortJobU32& q = jobs[j];
        in_tmp[j] = 0;
        if (q.n <= 1) continue;
        if (!radix_sort_takes_hist0(q.n)) {  // one workgroup, its own launch
            in_tmp[j] = radix_sort_impl<uint32_t>(q.keys, q.vals, q.tkeys, q.tvals, q.n, q.bits, q.hist, q.scan_tmp, s);
            continue;
        }
        r[nr] = SortRun<uint32_t>{q.keys, q.vals, q.tkeys, q.tvals, q.n, (q.n + SORT_TILE - 1) / SORT_TILE,
                                  (q.bits + 7) / 8, q.hist0_ready, q.hist, q.scan_tmp};
        map[nr++] = j;
    }
    if (nr == 0) return;
    int parity[2] = {0, 0};
    multi_tile_passes<uint32_t>(r, nr, s, parity);
    for (int k = 0; k < nr; ++k) in_tmp[map[k]] = parity[k];
}
int radix_sort_u64(uint64_t* keys, uint32_t* vals, uint64_t* tmp_keys, uint32_t* tmp_vals, uint32_t n, int bits,
                   uint32_t* hist, uint32_t* scan_tmp, hipStream_t s) {
    return radix_sort_impl<uint64_t>(keys, vals, tmp_keys, tmp_vals, n, bits, hist, scan_tmp, s);
}

// ------------------------------------------------- bucket (MSD) sorts

bool msd_sort_fits(uint32_t n, int bits) {
    return radix_sort_takes_hist0(n) && bits > kMsdBits && bits <= 2 * kMsdBits;
}

struct MsdSort {
    const uint32_t* keys; const uint32_t* vals;
    uint32_t* tk; uint32_t* tv;
    uint32_t* w;  // kMsdWords (layout in gi_device.h)
    uint32_t n;
    int shift;  // bits - kMsdBits: key >> shift is the bucket, the bits below it are the low part
    uint32_t slices;  // slice workgroups launched for the split buckets (a bound: unused ones exit)
};
struct MsdJobs {
    MsdSort s0, s1;
    uint32_t split;  // scatter blocks of sort 0
    uint32_t regular;  // bucket blocks: kMsdBuckets per sort
};

constexpr int MSD_THREADS = 256, MSD_ITEMS = 8, MSDB_ITEMS = 16;
constexpr uint32_t MSD_TILE = MSD_THREADS * MSD_ITEMS;
static_assert(MSD_THREADS == SCAN_THREADS, "block_exclusive_sum's block");
static_assert(kMsdBuckets == 16 * MSD_THREADS, "16 buckets per thread in the scans");
constexpr int SLICE_ITEMS = (int)(kMsdSlice / MSD_THREADS);  // a slice's keys per thread
static_assert(kMsdSlice % MSD_THREADS == 0, "whole items per thread in a slice");
static_assert(kMsdBigSlots < 255 && 33 <= kMsdOffDone - kMsdOffTab, "slot table");
constexpr uint32_t kMsdNoSlot = 0xFFFFFFFFu;

// Slice workgroups one sort needs at most: sum over split buckets of
// ceil(size / kMsdSlice) <= n / kMsdSlice + kMsdBigSlots.
uint32_t msd_slice_blocks(uint32_t n) { return blocks_for(n, kMsdSlice) + (uint32_t)kMsdBigSlots; }

// LDS arrays of bucket counts take one pad word per 32 entries: the prefix
// reads 16 consecutive entries per thread, a 16-word stride across the
// lanes, which unpadded falls on two of the 32 banks.
constexpr uint32_t kMsdLds = kMsdBuckets + kMsdBuckets / 32;
__device__ __forceinline__ uint32_t msd_pad(uint32_t i) { return i + (i >> 5); }

// Exclusive prefix of kMsdBuckets counts in LDS (padded, in place), 16
// consecutive ones per thread.
__device__ __forceinline__ void msd_prefix_lds(uint32_t* c) {
    constexpr int per = kMsdBuckets / MSD_THREADS;
    const uint32_t t0 = threadIdx.x * per;
    uint32_t v[per], sum = 0;
#pragma unroll
    for (int k = 0; k < per; ++k) {
        v[k] = c[msd_pad(t0 + k)];
        sum += v[k];
    }
    uint32_t off = block_exclusive_sum(sum, nullptr);
#pragma unroll
    for (int k = 0; k < per; ++k) {
        c[msd_pad(t0 + k)] = off;
        off += v[k];
    }
}

// 16 consecutive words from 16-byte aligned global memory
__device__ __forceinline__ void load16(const uint32_t* __restrict__ p, uint32_t (&v)[16]) {
    const uint4* __restrict__ p4 = (const uint4*)p;
#pragma unroll
    for (int q4 = 0; q4 < 4; ++q4) {
        const uint4 x = p4[q4];
        v[4 * q4] = x.x; v[4 * q4 + 1] = x.y; v[4 * q4 + 2] = x.z; v[4 * q4 + 3] = x.w;
    }
}

// Every key of a tile to its bucket: the buckets' starts (exclusive prefix
// of the totals, recomputed by every block) and the split buckets' slots
// (the first kMsdBigSlots buckets of more than kMsdBig keys, in bucket
// order; block 0 also stores both, and the slot table, for the bucket
// kernel), the tile's count per bucket with each key's rank among them (LDS
// atomics), one global reservation per bucket the tile holds, then the
// scatter into (tk, tv). Keys of split buckets also go into their slot's
// global low-part histogram, one atomic per run of equal keys in a wave.
__global__ __launch_bounds__(MSD_THREADS) void k_msd_scatter(MsdJobs J) {
    const bool second = blockIdx.x >= J.split;
    const MsdSort& S = second ? J.s1 : J.s0;
    const uint32_t tile = blockIdx.x - (second ? J.split : 0u);
    __shared__ uint32_t base[kMsdLds];
    __shared__ uint32_t cnt[kMsdLds];
    __shared__ __align__(16) uint8_t slot_of[kMsdBuckets];
    __shared__ uint32_t big_size[kMsdBigSlots];
    __shared__ uint32_t nbig_lds;
    const uint32_t t = threadIdx.x;
    {  // thread t: totals 16t .. 16t + 15 in four 16-byte loads, then the block's scans
        constexpr int per = kMsdBuckets / MSD_THREADS;
        uint32_t v[per], sum = 0, nb = 0;
        load16(S.w + per * t, v);
#pragma unroll
        for (int k = 0; k < per; ++k) {
            sum += v[k];
            nb += v[k] > kMsdBig ? 1u : 0u;
        }
        uint32_t slot, nbig;
        uint32_t off = block_exclusive_sum2(sum, nb, slot, nbig);
        uint32_t packed[per / 4] = {};  // the 16 slot bytes, one 16-byte LDS store
#pragma unroll
        for (int k = 0; k < per; ++k) {
            const uint32_t d = per * t + k;
            base[msd_pad(d)] = off;
            cnt[msd_pad(d)] = 0u;
            const bool big = v[k] > kMsdBig && slot < (uint32_t)kMsdBigSlots;
            packed[k / 4] |= (big ? slot : 0xFFu) << (8 * (k % 4));
            if (big) {
                big_size[slot] = v[k];
                if (tile == 0) S.w[kMsdOffTab + slot] = d;
            }
            slot += v[k] > kMsdBig ? 1u : 0u;
            off += v[k];
        }
        ((uint4*)slot_of)[t] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
        if (t == 0) nbig_lds = nbig < (uint32_t)kMsdBigSlots ? nbig : (uint32_t)kMsdBigSlots;
    }
    __syncthreads();
    if (tile == 0) {  // the starts and slots for the bucket kernel, coalesced
        for (uint32_t d = t; d < kMsdBuckets; d += MSD_THREADS) {
            S.w[kMsdOffStart + d] = base[msd_pad(d)];
            S.w[kMsdOffSlot + d] = slot_of[d] == 0xFFu ? kMsdNoSlot : (uint32_t)slot_of[d];
        }
        if (t < 64) {  // first slice of each slot: 33 entries, the last repeated past the used slots
            const uint32_t c = t < nbig_lds ? (big_size[t] + kMsdSlice - 1u) / kMsdSlice : 0u;
            uint32_t x = c;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(x, o);
                if (t >= (uint32_t)o) x += y;
            }
            if (t <= (uint32_t)kMsdBigSlots) S.w[kMsdOffTab + kMsdBigSlots + t] = x - c;
        }
    }
    const uint32_t mask = (1u << S.shift) - 1u;
    uint32_t key[MSD_ITEMS], val[MSD_ITEMS], rank[MSD_ITEMS];
#pragma unroll
    for (int it = 0; it < MSD_ITEMS; ++it) {
        const uint32_t i = tile * MSD_TILE + it * MSD_THREADS + t;
        key[it] = i < S.n ? S.keys[i] : 0u;
        val[it] = i < S.n ? S.vals[i] : 0u;
    }
#pragma unroll
    for (int it = 0; it < MSD_ITEMS; ++it) {
        const uint32_t d = key[it] >> S.shift;
        const WaveRun r = wave_run(d, tile * MSD_TILE + it * MSD_THREADS + t < S.n);
        rank[it] = wave_run_place(r.len ? atomicAdd(&cnt[msd_pad(d)], r.len) : 0u, r);
    }
    __syncthreads();
    for (uint32_t d = t; d < kMsdBuckets; d += MSD_THREADS) {
        const uint32_t c = cnt[msd_pad(d)];
        if (c) cnt[msd_pad(d)] = atomicAdd(&S.w[kMsdOffCursor + d], c);  // the tile's place in the bucket
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < MSD_ITEMS; ++it) {
        if (tile * MSD_TILE + it * MSD_THREADS + t >= S.n) continue;
        const uint32_t d = msd_pad(key[it] >> S.shift);
        const uint32_t pos = base[d] + cnt[d] + rank[it];
        if (pos < S.n) {  // always, when the totals count these keys
            S.tk[pos] = key[it];
            S.tv[pos] = val[it];
        }
    }
    // the split buckets' histograms last, where no barrier waits for these
    // atomics; waves without such keys skip them after one ballot
    uint32_t bigbits = 0;
#pragma unroll
    for (int it = 0; it < MSD_ITEMS; ++it)
        if (tile * MSD_TILE + it * MSD_THREADS + t < S.n && slot_of[key[it] >> S.shift] != 0xFFu) bigbits |= 1u << it;
    if (__ballot(bigbits != 0u)) {
#pragma unroll
        for (int it = 0; it < MSD_ITEMS; ++it) {
            const bool big = (bigbits >> it) & 1u;
            if (__ballot(big)) {
                const WaveRun rb = wave_run(key[it], big);
                if (rb.len)
                    atomicAdd(&S.w[kMsdOffHist + slot_of[key[it] >> S.shift] * kMsdBuckets + (key[it] & mask)], rb.len);
            }
        }
    }
}

// One slice of a split bucket: its kMsdSlice keys counted per low part in
// LDS, each low part's places reserved with one global atomic on the slot's
// cursor (after the bucket's prefix from the slot's histogram), then every
// key to its place. The last slice of the bucket to finish clears the slot's
// histogram, cursors and done counter for the next sort.
__device__ void msd_slice(const MsdSort& S, uint32_t e, uint32_t* cnt) {
    const uint32_t t = threadIdx.x;
    const uint32_t* tab = S.w + kMsdOffTab;
    const uint32_t* first = tab + kMsdBigSlots;
    if (e >= first[kMsdBigSlots]) return;  // no such slice in this sort
    uint32_t slot = 0;
    for (uint32_t k = 1; k < (uint32_t)kMsdBigSlots; ++k) slot += first[k] <= e ? 1u : 0u;
    const uint32_t g = e - first[slot], b = tab[slot];
    const uint32_t size = S.w[b], start = S.w[kMsdOffStart + b];
    const uint32_t nslices = first[slot + 1] - first[slot];
    uint32_t* hist = S.w + kMsdOffHist + slot * kMsdBuckets;
    uint32_t* cur = S.w + kMsdOffBigCursor + slot * kMsdBuckets;
    if (start + size <= S.n) {  // (always, when the totals count the keys)
        uint32_t* keys = (uint32_t*)S.keys;
        uint32_t* vals = (uint32_t*)S.vals;
        const uint32_t mask = (1u << S.shift) - 1u;
        for (uint32_t d = t; d < kMsdBuckets; d += MSD_THREADS) cnt[msd_pad(d)] = 0u;
        __syncthreads();
        const uint32_t j0 = g * kMsdSlice, j1 = j0 + kMsdSlice < size ? j0 + kMsdSlice : size;
        uint32_t k[SLICE_ITEMS], v[SLICE_ITEMS];
#pragma unroll
        for (int it = 0; it < SLICE_ITEMS; ++it) {
            const uint32_t j = j0 + it * MSD_THREADS + t;
            k[it] = j < j1 ? S.tk[start + j] : 0u;
            v[it] = j < j1 ? S.tv[start + j] : 0u;
        }
#pragma unroll
        for (int it = 0; it < SLICE_ITEMS; ++it)
            if (j0 + it * MSD_THREADS + t < j1) atomicAdd(&cnt[msd_pad(k[it] & mask)], 1u);
        __syncthreads();
        {  // thread t: low parts 16t .. 16t + 15 of the bucket's histogram
            uint32_t h[16], sum = 0;
            load16(hist + 16 * t, h);
#pragma unroll
            for (int q = 0; q < 16; ++q) sum += h[q];
            uint32_t off = start + block_exclusive_sum(sum, nullptr);
            uint32_t own[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) own[q] = cnt[msd_pad(16 * t + q)];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                if (own[q]) cnt[msd_pad(16 * t + q)] = off + atomicAdd(&cur[16 * t + q], own[q]);
                off += h[q];
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < SLICE_ITEMS; ++it)
            if (j0 + it * MSD_THREADS + t < j1) {
                const uint32_t p = atomicAdd(&cnt[msd_pad(k[it] & mask)], 1u);
                if (p < S.n) {
                    keys[p] = k[it];
                    vals[p] = v[it];
                }
            }
    }
    // the last slice of the bucket clears its slot. No device-scope fence:
    // the barrier waits for this block's loads and atomics, whose results
    // are all consumed before it, and the clearing stores depend on the
    // done count; they reach the next sort through the launch boundary. (A
    // fence here writes back and invalidates the XCD's L2, which slowed
    // every workgroup of the launch threefold.)
    __shared__ uint32_t last;
    __syncthreads();
    if (t == 0) last = atomicAdd(&S.w[kMsdOffDone + slot], 1u) == nslices - 1u ? 1u : 0u;
    __syncthreads();
    if (last) {
        uint4* h4 = (uint4*)hist;
        uint4* c4 = (uint4*)cur;
        for (uint32_t i = t; i < kMsdBuckets / 4; i += MSD_THREADS) {
            h4[i] = make_uint4(0u, 0u, 0u, 0u);
            c4[i] = make_uint4(0u, 0u, 0u, 0u);
        }
        if (t == 0) atomicExch(&S.w[kMsdOffDone + slot], 0u);
    }
}

// One bucket per workgroup, sorted by its low bits (at most kMsdBits):
// counts, their prefix, then every key to its place (an atomic on the count
// of its low part), from (tk, tv) back into (keys, vals); the blocks before
// the regular ones take the slices of the split buckets (msd_slice), whose
// regular blocks exit. Measured against: the counts and places in runs of
// equal low parts (one atomic per run), 36.5 -> 46.7 us, since each item
// then waits on its atomic's result; and a fixed grid taking the non-empty
// buckets from a list, 36.5 -> 40.8 us.
__global__ __launch_bounds__(MSD_THREADS) void k_msd_buckets(MsdJobs J) {
    __shared__ uint32_t cnt[kMsdLds];
    const uint32_t slices = J.s0.slices + J.s1.slices;
    if (blockIdx.x < slices) {  // first, as the longest workgroups
        if (blockIdx.x < J.s0.slices) msd_slice(J.s0, blockIdx.x, cnt);
        else msd_slice(J.s1, blockIdx.x - J.s0.slices, cnt);
        return;
    }
    const uint32_t r = blockIdx.x - slices;
    const bool second = r >= kMsdBuckets;
    const MsdSort& S = second ? J.s1 : J.s0;
    const uint32_t b = r - (second ? kMsdBuckets : 0u);
    const uint32_t size = S.w[b];
    const uint32_t start = S.w[kMsdOffStart + b], t = threadIdx.x;
    if (size == 0 || start + size > S.n) return;  // (the second never, when the totals count the keys)
    if (size > kMsdBig && S.w[kMsdOffSlot + b] != kMsdNoSlot) return;  // split: the slice blocks sort it
    uint32_t* keys = (uint32_t*)S.keys;
    uint32_t* vals = (uint32_t*)S.vals;
    if (size == 1) {
        if (t == 0) {
            keys[start] = S.tk[start];
            vals[start] = S.tv[start];
        }
        return;
    }
    const uint32_t mask = (1u << S.shift) - 1u;
    for (uint32_t d = t; d < kMsdBuckets; d += MSD_THREADS) cnt[msd_pad(d)] = 0u;
    __syncthreads();
    // rounds of up to MSDB_ITEMS keys per thread, all loaded before any is
    // counted (a bucket's items end at the same item number on every thread)
    for (uint32_t r0 = 0; r0 < size; r0 += MSDB_ITEMS * MSD_THREADS) {
        uint32_t k[MSDB_ITEMS];
#pragma unroll
        for (int it = 0; it < MSDB_ITEMS; ++it) {
            if (r0 + it * MSD_THREADS >= size) break;
            const uint32_t j = r0 + it * MSD_THREADS + t;
            k[it] = j < size ? S.tk[start + j] : 0u;
        }
#pragma unroll
        for (int it = 0; it < MSDB_ITEMS; ++it) {
            if (r0 + it * MSD_THREADS >= size) break;
            if (r0 + it * MSD_THREADS + t < size) atomicAdd(&cnt[msd_pad(k[it] & mask)], 1u);
        }
    }
    __syncthreads();
    msd_prefix_lds(cnt);
    __syncthreads();
    for (uint32_t r0 = 0; r0 < size; r0 += MSDB_ITEMS * MSD_THREADS) {
        uint32_t k[MSDB_ITEMS], v[MSDB_ITEMS];
#pragma unroll
        for (int it = 0; it < MSDB_ITEMS; ++it) {
            if (r0 + it * MSD_THREADS >= size) break;
            const uint32_t j = r0 + it * MSD_THREADS + t;
            k[it] = j < size ? S.tk[start + j] : 0u;
            v[it] = j < size ? S.tv[start + j] : 0u;
        }
#pragma unroll
        for (int it = 0; it < MSDB_ITEMS; ++it) {
            if (r0 + it * MSD_THREADS >= size) break;
            if (r0 + it * MSD_THREADS + t < size) {
                const uint32_t p = atomicAdd(&cnt[msd_pad(k[it] & mask)], 1u);
                keys[start + p] = k[it];
                vals[start + p] = v[it];
            }
        }
    }
}

void msd_sort_pair(const MsdSortJob* jobs, int njobs, hipStream_t s) {
    if (njobs <= 0) return;
    MsdJobs J;
    MsdSort* S[2] = {&J.s0, &J.s1};
    for (int k = 0; k < 2; ++k) {
        const MsdSortJob& q = jobs[k < njobs ? k : 0];
        *S[k] = MsdSort{q.keys, q.vals, q.tkeys, q.tvals, q.words, q.n, q.bits - kMsdBits,
                        k < njobs ? msd_slice_blocks(q.n) : 0u};
    }
    J.split = blocks_for(jobs[0].n, MSD_TILE);
    J.regular = (njobs > 1 ? 2u : 1u) * kMsdBuckets;
    const uint32_t tiles = J.split + (njobs > 1 ? blocks_for(jobs[1].n, MSD_TILE) : 0u);
    k_msd_scatter<<<tiles, MSD_THREADS, 0, s>>>(J);
    k_msd_buckets<<<J.regular + J.s0.slices + J.s1.slices, MSD_THREADS, 0, s>>>(J);
    GI_LAUNCH_CHECK();
}

// --------------------------------------------------------- BVH (Karras)
//
// Build and refit are one kernel per chunk of CHUNK (256 or 512) sorted
// leaves, one thread per leaf and per internal node: the chunk keeps in LDS
// its keys' adjacent common-prefix lengths and their window minima, from
// which each internal node whose leaf range lies inside the chunk gets its
// range and split by binary lifting (karras_table_node: about 20 byte reads
// instead of Karras' two searches, whose ~1,000 VALU instructions per node
// made this kernel VALU-bound), and takes the union of its leaves from a
// table of aligned block boxes (1, 2, 4, ... 64 leaves) built in LDS by wave
// butterflies. No bottom-up merge: no parents, flags or atomics (a
// global-atomic bottom-up refit, with the device-scope fences it needs on
// gfx950's per-XCD L2s, cost 6 ms for 2^20 triangles; an LDS-atomic one
// still a quarter of this kernel). The few "top" nodes whose range crosses a
// chunk boundary keep the box of their leaves up to the chunk edge; their
// searches over the whole key array (whose dependent key loads held a wave
// for ~10 us here) and the rest of their box, from a 64-ary hierarchy of
// block boxes, are k_top_boxes' work.
// A refit after a rigid motion reruns the same kernels on the same sorted
// keys: the topology comes out identical, only the boxes change.

// Up to two trees per launch: blocks [0, split) build tree 0, the rest tree 1.
// Fields are picked with selects on a block-uniform flag, so the kernel
// arguments are never indexed dynamically.
template <class K>
struct ChunkJobs {
    const K* keys0; const K* keys1;
    const uint32_t* prims0; const uint32_t* prims1;
    const float4* box0; const float4* box1;  // triangle boxes, (lo, hi) per triangle
    float4* nodes0; float4* nodes1;
    uint2* range0; uint2* range1;
    float4* level1_0; float4* level1_1;  // boxes of the 64-leaf blocks (first level of the block hierarchy)
    uint64_t* topmask0; uint64_t* topmask1;  // per 64 internal nodes: which are top nodes
    int n0, n1;
    uint32_t split;
};

// First box of level k in a chunk's aligned-block table (level k holds the
// CHUNK >> k boxes of leaves [m 2^k, (m + 1) 2^k), k = 0..6).
__host__ __device__ constexpr int block_table_offset(int chunk, int k) { return 2 * (chunk - (chunk >> k)); }

template <class K, int CHUNK>
__global__ __launch_bounds__(CHUNK) void k_build_chunks(ChunkJobs<K> J) {
    bool second;
    const uint32_t chunk = xcd_block_split(blockIdx.x, gridDim.x, J.split, second);
    const K* __restrict__ keys = second ? J.keys1 : J.keys0;
    const uint32_t* __restrict__ prims = second ? J.prims1 : J.prims0;
    const float4* __restrict__ tbox = second ? J.box1 : J.box0;
    float4* __restrict__ nodes = second ? J.nodes1 : J.nodes0;
    uint2* __restrict__ range = second ? J.range1 : J.range0;
    float4* __restrict__ level1 = second ? J.level1_1 : J.level1_0;
    uint64_t* __restrict__ topmask = second ? J.topmask1 : J.topmask0;
    const int n = second ? J.n1 : J.n0;
    constexpr int kLevels = DeltaLevels<CHUNK>::kLevels;
    __shared__ K skey[CHUNK];
    __shared__ float4 sbox[2 * block_table_offset(CHUNK, 7)];  // (lo, hi) pairs of levels 0..6
    __shared__ int8_t dtab[kLevels][CHUNK + 1];                // adjacent prefixes and their window minima
    const int t = (int)threadIdx.x;
    const int cs = (int)chunk * CHUNK;
    const int ce = min(n, cs + CHUNK);
    const int jl = cs + t;  // this thread's leaf, and its internal node
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    uint32_t prim = 0;
    {  // A[jl] = delta(jl, jl + 1) at position t + 1; thread 0 also A[cs - 1] at 0
        int a = -1;
        if (jl + 1 < ce || (jl + 1 == ce && ce < n)) a = delta_keys(keys[jl], jl, keys[jl + 1], jl + 1);
        dtab[0][t + 1] = (int8_t)a;
        if (t == 0) dtab[0][0] = (int8_t)(cs > 0 ? delta_keys(keys[cs - 1], cs - 1, keys[cs], cs) : -1);
    }
    if (jl < ce) {
        skey[t] = keys[jl];
        prim = prims[jl];
        float4 l = tbox[2 * (size_t)prim], h = tbox[2 * (size_t)prim + 1];
        b[0] = l.x; b[1] = l.y; b[2] = l.z;
        b[3] = h.x; b[4] = h.y; b[5] = h.z;
    }
    const float leaf[6] = {b[0], b[1], b[2], b[3], b[4], b[5]};
    sbox[2 * t] = make_float4(b[0], b[1], b[2], 0.f);
    sbox[2 * t + 1] = make_float4(b[3], b[4], b[5], 0.f);
    // aligned 2^k-leaf blocks by butterfly reduction within the wave (all
    // lanes active): after the exchange at distance 2^(k-1) a lane holds the
    // union of its block
    auto level_box = [&](int k) {
        if ((t & ((1 << k) - 1)) == 0) {
            const int e = 2 * (block_table_offset(CHUNK, k) + (t >> k));
            sbox[e] = make_float4(b[0], b[1], b[2], 0.f);
            sbox[e + 1] = make_float4(b[3], b[4], b[5], 0.f);
        }
    };
    box_xor_step<1>(b);
    level_box(1);
    box_xor_step<2>(b);
    level_box(2);
    box_xor_step<4>(b);
    level_box(3);
    box_xor_step<8>(b);
    level_box(4);
    box_xor_step<16>(b);
    level_box(5);
    box_xor_step<32>(b);
    level_box(6);
    // the 64-leaf block is the first level of the block hierarchy
    if ((t & 63) == 0 && jl < ce) {
        level1[2 * (size_t)(jl >> 6)] = make_float4(b[0], b[1], b[2], 0.f);
        level1[2 * (size_t)(jl >> 6) + 1] = make_float4(b[3], b[4], b[5], 0.f);
    }
    __syncthreads();
    // window minima of the adjacent prefixes, level by level (windows that
    // pass position CHUNK are never read)
#pragma unroll
    for (int L = 1; L < kLevels; ++L) {
        if (t + (1 << L) - 1 <= CHUNK) {
            const int8_t u = dtab[L - 1][t], v = dtab[L - 1][t + (1 << (L - 1))];
            dtab[L][t] = u < v ? u : v;
        }
        __syncthreads();
    }
    const KeySrc<K> KS{keys, skey, cs, ce, n};
    if (jl < ce) {
        const uint32_t rope = jl == n - 1 ? kSentinel : chain_top(KS, jl + 1);
        nodes[2 * (size_t)(n - 1 + jl)] = make_float4(leaf[0], leaf[1], leaf[2], __uint_as_float(prim));
        nodes[2 * (size_t)(n - 1 + jl) + 1] = make_float4(leaf[3], leaf[4], leaf[5], __uint_as_float(rope));
    }
    if (jl >= min(ce, n - 1)) return;
    int lo_i, hi_i;
    uint32_t left = 0, right = 0, rope = 0;
    // in-chunk nodes from the table; a top node, whose range leaves the
    // chunk on its side d, keeps here the box of its leaves up to the chunk
    // edge and gets its range, children and rope from k_top_boxes
    const bool top = !karras_table_node<CHUNK, K>(KS, dtab, jl, lo_i, hi_i, left, right, rope);
    if (top) {
        const bool up = dtab[0][t + 1] > dtab[0][t];  // d = +1: delta(i, i + 1) > delta(i, i - 1)
        lo_i = up ? jl : cs;
        hi_i = up ? ce - 1 : jl;
    } else {
        range[jl] = make_uint2((uint32_t)lo_i, (uint32_t)hi_i);
    }
    // which of the wave's 64 internal nodes are top nodes (lane 0 holds the
    // first; a wave with none left has no internal node)
    const uint64_t tops = __ballot(top);
    if ((t & 63) == 0) topmask[jl >> 6] = tops;
    // the union of the range's leaves inside the chunk (all of them for an
    // in-chunk node; k_top_boxes adds a top node's leaves beyond the chunk
    // edge) as disjoint aligned blocks, at most two per level below 64 and
    // whole 64-leaf blocks above
    float q[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    int L = max(lo_i, cs) - cs, R = min(hi_i, ce - 1) - cs + 1;
    auto absorb = [&](int k, int m) {
        const int e = 2 * (block_table_offset(CHUNK, k) + m);
        const float4 l = sbox[e], h = sbox[e + 1];
        q[0] = fminf(q[0], l.x); q[1] = fminf(q[1], l.y); q[2] = fminf(q[2], l.z);
        q[3] = fmaxf(q[3], h.x); q[4] = fmaxf(q[4], h.y); q[5] = fmaxf(q[5], h.z);
    };
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (L < R && ((L >> k) & 1)) { absorb(k, L >> k); L += 1 << k; }
        if (L < R && ((R >> k) & 1)) { absorb(k, (R >> k) - 1); R -= 1 << k; }
    }
    for (int m = L >> 6; m < (R >> 6); ++m) absorb(6, m);
    const uint32_t lw = left | (right >= (uint32_t)(n - 1) ? kRightLeaf : 0u);  // top nodes: k_top_boxes
    nodes[2 * (size_t)jl] = make_float4(q[0], q[1], q[2], __uint_as_float(lw));
    nodes[2 * (size_t)jl + 1] = make_float4(q[3], q[4], q[5], __uint_as_float(rope));
}

// one wave per output block of 64 input elements; blocks [0, split) serve
// tree 0, the rest tree 1
struct BlockBoxJobs {
    const float4* in0; const float4* in1;
    float4* out0; float4* out1;
    uint32_t n_in0, n_in1, n_out0, n_out1;
    uint32_t split;
};

__global__ __launch_bounds__(256) void k_block_boxes(BlockBoxJobs J) {
    const bool second = blockIdx.x >= J.split;
    const float4* __restrict__ in = second ? J.in1 : J.in0;
    float4* __restrict__ out = second ? J.out1 : J.out0;
    const uint32_t n_in = second ? J.n_in1 : J.n_in0, n_out = second ? J.n_out1 : J.n_out0;
    uint32_t blk = (blockIdx.x - (second ? J.split : 0u)) * 4 + (threadIdx.x >> 6);
    if (blk >= n_out) return;
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    uint32_t idx = blk * 64 + lane_id();
    if (idx < n_in) box_absorb(in, idx, b);
    wave_box_reduce(b);
    if (lane_id() == 0) {
        out[2 * (size_t)blk] = make_float4(b[0], b[1], b[2], 0.f);
        out[2 * (size_t)blk + 1] = make_float4(b[3], b[4], b[5], 0.f);
    }
}

// The top nodes: range, children and rope, then the box. The chunk kernel
// left in each the box of its leaves up to its chunk's edge (on the side
// its range leaves the chunk). Karras' two searches run over the whole key
// array, 16 probes per round on a quarter-wave's lanes (karras_top_node16);
// the leaves beyond the chunk edge then come from the block hierarchy as
// disjoint aligned windows, at most two per level. A wave takes TOP_GROUPS
// groups of 64 internal nodes (the chunk kernel's masks) and gives each of
// their top nodes a quarter of the wave, four nodes at a time, so that the
// nodes' load chains overlap; a quarter issues all its window loads before
// it reduces. Blocks [0, split) serve tree 0.
constexpr uint32_t TOP_GROUPS = 4, TOP_LANES = 16;
constexpr uint32_t kTopLevelBoxes = 256;  // boxes of the highest block level k_top_boxes reads whole

struct TopBoxJobs {
    BlockLevels lv0, lv1;
    const uint32_t* keys0; const uint32_t* keys1;  // the sorted keys
    uint2* range0; uint2* range1;
    const uint64_t* topmask0; const uint64_t* topmask1;
    float4* nodes0; float4* nodes1;
    uint32_t n0, n1;
    int top0, top1;  // the highest block level built (at most kTopLevelBoxes boxes)
    uint32_t split;
};

__global__ __launch_bounds__(256) void k_top_boxes(TopBoxJobs J, uint32_t chunk) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const bool second = blk >= J.split;
    const uint32_t n = second ? J.n1 : J.n0;
    uint2* __restrict__ range = second ? J.range1 : J.range0;
    const uint32_t* __restrict__ keys = second ? J.keys1 : J.keys0;
    float4* nodes = second ? J.nodes1 : J.nodes0;
    const BlockLevels& lv = second ? J.lv1 : J.lv0;  // read in place from the kernel arguments
    const int top = second ? J.top1 : J.top0;
    const uint32_t groups = n > 1 ? (n - 1 + 63) / 64 : 0u;  // groups of the n - 1 internal nodes
    const uint32_t g0 = ((blk - (second ? J.split : 0u)) * 4 + (threadIdx.x >> 6)) * TOP_GROUPS;
    if (g0 >= groups) return;
    const uint32_t lane = lane_id(), quarter = lane / TOP_LANES, ql = lane % TOP_LANES;
    const uint64_t* __restrict__ topmask = second ? J.topmask1 : J.topmask0;
    const uint64_t own = lane < TOP_GROUPS && g0 + lane < groups ? topmask[g0 + lane] : 0ull;
    uint64_t m[TOP_GROUPS];
    uint32_t c[TOP_GROUPS], total = 0;
#pragma unroll
    for (uint32_t g = 0; g < TOP_GROUPS; ++g) {
        m[g] = __shfl(own, (int)g);
        c[g] = (uint32_t)__popcll(m[g]);
        total += c[g];
    }
    for (uint32_t k = quarter; k < total; k += 64 / TOP_LANES) {
        // the k-th top node of the wave's groups
        uint32_t g = 0, kk = k;
        uint64_t mm = m[0];
#pragma unroll
        for (uint32_t q = 0; q + 1 < TOP_GROUPS; ++q)
            if (g == q && kk >= c[q]) {
                kk -= c[q];
                g = q + 1;
                mm = m[q + 1];
            }
        for (; kk > 0; --kk) mm &= mm - 1;
        const uint32_t node = (g0 + g) * 64 + (uint32_t)__ffsll((long long)mm) - 1;
        const float4 part_lo = nodes[2 * (size_t)node], part_hi = nodes[2 * (size_t)node + 1];
        const uint32_t cs = node / chunk * chunk, ce = min(cs + chunk, n);
        // range, children and rope: Karras' searches over the whole key
        // array, 16 probes per round on the quarter's lanes. The rope's keys
        // and the box windows need only the range: they are loaded before
        // the split search, off its chain of dependent rounds.
        const int i = (int)node, nn = (int)n;
        const uint32_t ki = keys[i];
        const int dnext = delta_keys(ki, i, keys[i + 1], i + 1);
        const int dprev = i > 0 ? delta_keys(ki, i, keys[i - 1], i - 1) : -1;
        const int d = dnext > dprev ? 1 : -1, dmin = d > 0 ? dprev : dnext;
        const KeySrc<uint32_t> KS{keys, nullptr, 0, 0, nn};
        auto count = [&](auto pos, auto pred) {
            const bool ok = pred(pos((int)ql));
            return (int)__popcll((__ballot(ok) >> (TOP_LANES * quarter)) & 0xFFFFull);
        };
        const int l = karras_top_range16(KS, i, ki, d, dmin, d > 0 ? (int)ce - i : i - (int)cs + 1, count);
        const int j = i + l * d;
        const uint2 r = make_uint2((uint32_t)min(i, j), (uint32_t)max(i, j));
        if (ql == 0) range[node] = r;
        // the keys the split search starts from and the rope needs, loaded
        // together before the windows
        const uint32_t kj = keys[j], k_hi = keys[r.y];
        const uint32_t k_hi1 = r.y + 1 < n ? keys[r.y + 1] : 0u, k_hi2 = r.y + 2 < n ? keys[r.y + 2] : 0u;
        uint32_t L = r.x < cs ? r.x : ce, R = r.x < cs ? cs : r.y + 1;  // the leaves [L, R) beyond the chunk
        // windows (first element, count) per level and side
        uint32_t wS[2 * kMaxBlockLevels], wC[2 * kMaxBlockLevels];
        bool done = false;
#pragma unroll
        for (int lvl = 0; lvl < kMaxBlockLevels; ++lvl) {
            wS[2 * lvl] = wS[2 * lvl + 1] = 0;
            wC[2 * lvl] = wC[2 * lvl + 1] = 0;
            if (done) continue;
            if (R - L <= 64 || lvl == top) {  // the whole rest at this level (at most kTopLevelBoxes at the top)
                wS[2 * lvl] = L;
                wC[2 * lvl] = R - L;
                done = true;
                continue;
            }
            const uint32_t Lh = (L + 63u) & ~63u, Rt = R & ~63u;
            wS[2 * lvl] = L;
            wC[2 * lvl] = Lh - L;
            wS[2 * lvl + 1] = Rt;
            wC[2 * lvl + 1] = R - Rt;
            L = Lh >> 6;
            R = Rt >> 6;
            if (L >= R) done = true;
        }
        float b[6] = {part_lo.x, part_lo.y, part_lo.z, part_hi.x, part_hi.y, part_hi.z};
#pragma unroll
        for (int w = 0; w < 2 * kMaxBlockLevels; ++w) {
#pragma unroll
            for (uint32_t e = 0; e < 64; e += TOP_LANES)
                if (ql + e < wC[w]) box_absorb(lv.level[w / 2], wS[w] + ql + e, b);
            for (uint32_t e = 64; e < wC[w]; e += TOP_LANES)  // a top-level window beyond 64 boxes
                if (ql + e < wC[w]) box_absorb(lv.level[w / 2], wS[w] + ql + e, b);
        }
#pragma unroll
        for (int o = TOP_LANES / 2; o > 0; o >>= 1) {  // within the quarter
#pragma unroll
            for (int q = 0; q < 3; ++q) b[q] = fminf(b[q], __shfl_xor(b[q], o));
#pragma unroll
            for (int q = 3; q < 6; ++q) b[q] = fmaxf(b[q], __shfl_xor(b[q], o));
        }
        const uint32_t rope = r.y == n - 1 ? kSentinel : chain_top_of(nn, (int)r.y + 1, k_hi, k_hi1, k_hi2);
        uint32_t left, right;
        karras_children(nn, (int)r.x, (int)r.y, karras_top_split16(KS, i, ki, d, l, delta_keys(ki, i, kj, j), count),
                        left, right);
        const uint32_t lw = left | (right >= n - 1 ? kRightLeaf : 0u);
        if (ql == 0) {  // .w: the left child (right-leaf flag) and the rope
            nodes[2 * (size_t)node] = make_float4(b[0], b[1], b[2], __uint_as_float(lw));
            nodes[2 * (size_t)node + 1] = make_float4(b[3], b[4], b[5], __uint_as_float(rope));
        }
    }
}

uint32_t block_level_count(uint32_t n, int level) {
    uint64_t c = n;
    for (int k = 0; k < level; ++k) c = (c + 63) / 64;
    return (uint32_t)c;
}

// float4 offset of the top-node masks (one uint64 per 64 internal nodes)
// after the block levels
static size_t top_mask_offset(uint32_t n) {
    size_t total = 0;
    for (int k = 1; k < kMaxBlockLevels; ++k) total += 2 * (size_t)block_level_count(n, k);
    return total;
}

size_t block_levels_elems(uint32_t n) { return top_mask_offset(n) + (block_level_count(n, 1) + 1) / 2; }

template <class K>
static void build_bvh_impl(const BuildJob<K>* jobs, int njobs, uint32_t chunk, hipStream_t s, hipEvent_t chunk_begin,
                           hipEvent_t chunk_end, unsigned event_flags) {
    const BuildJob<K>* j0 = &jobs[0];
    const BuildJob<K>* j1 = njobs > 1 ? &jobs[1] : nullptr;
    const uint32_t n0 = j0->n, n1 = j1 ? j1->n : 0;
    if (n0 + n1 == 0) return;
    ChunkJobs<K> cj;
    cj.keys0 = j0->keys; cj.prims0 = j0->prims; cj.box0 = j0->boxes;
    cj.nodes0 = j0->nodes; cj.range0 = j0->range; cj.level1_0 = j0->blocks; cj.n0 = (int)n0;
    const BuildJob<K>* o = j1 ? j1 : j0;
    cj.keys1 = o->keys; cj.prims1 = o->prims; cj.box1 = o->boxes;
    cj.nodes1 = o->nodes; cj.range1 = o->range; cj.level1_1 = o->blocks; cj.n1 = (int)n1;
    uint64_t* topmask[2] = {(uint64_t*)(j0->blocks + top_mask_offset(n0)),
                            (uint64_t*)(o->blocks + top_mask_offset(n1))};
    cj.topmask0 = topmask[0];
    cj.topmask1 = topmask[1];
    cj.split = blocks_for(n0, chunk);
    const uint32_t chunk_blocks = cj.split + blocks_for(n1, chunk);
    if (chunk_begin) (void)hipEventRecordWithFlags(chunk_begin, s, event_flags);
    if (chunk == 256) k_build_chunks<K, 256><<<chunk_blocks, 256, 0, s>>>(cj);
    else k_build_chunks<K, 512><<<chunk_blocks, 512, 0, s>>>(cj);
    if (chunk_end) (void)hipEventRecordWithFlags(chunk_end, s, event_flags);
    // 64-ary levels of leaf boxes for the nodes that cross chunks
    BlockLevels lv[2];
    const BuildJob<K>* jj[2] = {j0, o};
    uint32_t nn[2] = {n0, n1};
    for (int t = 0; t < 2; ++t) {
        lv[t].level[0] = jj[t]->nodes + 2 * (size_t)(nn[t] ? nn[t] - 1 : 0);
        size_t off = 0;
        for (int k = 1; k < kMaxBlockLevels; ++k) {
            lv[t].level[k] = jj[t]->blocks + off;
            off += 2 * (size_t)block_level_count(nn[t], k);
        }
    }
    // level 1 (64-leaf blocks) was written by k_build_chunks. Level k is built
    // only when level k - 1 has more than kTopLevelBoxes boxes: k_top_boxes
    // takes the rest of a range at the highest level built whole (2^20 leaves:
    // levels 1 and 2, 16,384 and 256 boxes, one launch).
    int top[2] = {1, 1};
    for (int k = 2; k < kMaxBlockLevels; ++k) {
        BlockBoxJobs bj;
        uint32_t nin[2], nout[2];
        for (int t = 0; t < 2; ++t) {
            nin[t] = block_level_count(nn[t], k - 1);
            nout[t] = nin[t] > kTopLevelBoxes ? block_level_count(nn[t], k) : 0;
            if (nout[t]) top[t] = k;
        }
        if (nout[0] == 0 && nout[1] == 0) break;
        bj.in0 = lv[0].level[k - 1]; bj.out0 = (float4*)lv[0].level[k]; bj.n_in0 = nin[0]; bj.n_out0 = nout[0];
        bj.in1 = lv[1].level[k - 1]; bj.out1 = (float4*)lv[1].level[k]; bj.n_in1 = nin[1]; bj.n_out1 = nout[1];
        bj.split = blocks_for(nout[0], 4);
        k_block_boxes<<<bj.split + blocks_for(nout[1], 4), 256, 0, s>>>(bj);
    }
    static_assert(sizeof(K) == 4, "top-node searches read 32-bit keys");
    TopBoxJobs tj;
    tj.lv0 = lv[0]; tj.lv1 = lv[1];
    tj.keys0 = (const uint32_t*)j0->keys; tj.keys1 = (const uint32_t*)o->keys;
    tj.range0 = j0->range; tj.range1 = o->range;
    tj.topmask0 = topmask[0]; tj.topmask1 = topmask[1];
    tj.nodes0 = j0->nodes; tj.nodes1 = o->nodes;
    tj.n0 = nn[0]; tj.n1 = nn[1];
    tj.top0 = top[0]; tj.top1 = top[1];
    // one wave per TOP_GROUPS groups of 64 internal nodes (n - 1 of them); a
    // one-leaf tree has none
    auto top_waves = [](uint32_t m) { return m > 1 ? blocks_for(blocks_for(m - 1, 64), TOP_GROUPS) : 0u; };
    tj.split = blocks_for(top_waves(nn[0]), 4);
    uint32_t top_blocks = tj.split + blocks_for(top_waves(nn[1]), 4);
    if (top_blocks) k_top_boxes<<<top_blocks, 256, 0, s>>>(tj, chunk);
    GI_LAUNCH_CHECK();
}

void launch_build_bvh(const BuildJob<uint32_t>* jobs, int njobs, uint32_t chunk, hipStream_t s,
                      hipEvent_t chunk_begin, hipEvent_t chunk_end, unsigned event_flags) {
    build_bvh_impl<uint32_t>(jobs, njobs, chunk, s, chunk_begin, chunk_end, event_flags);
}

// Subtree of each forest mesh: the node whose leaf range is exactly the
// mesh's [toff[m], toff[m+1]) (in Karras' numbering internal node first or
// internal node last; a one-triangle mesh is its leaf), and the rope that
// leaves it, which ends a traversal of that mesh. Empty mesh: both sentinel.
__global__ void k_forest_roots(const uint2* __restrict__ range, const float4* __restrict__ nodes, uint32_t n,
                               const uint32_t* __restrict__ toff, uint32_t M, uint2* __restrict__ roots) {
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    uint32_t f = toff[m], e = toff[m + 1];
    if (e == f) {
        roots[m] = make_uint2(kSentinel, kSentinel);
        return;
    }
    uint32_t l = e - 1, root;
    if (e - f == 1) {
        root = n - 1 + f;
    } else {
        uint2 r = range[f];
        root = (r.x == f && r.y == l) ? f : l;
    }
    roots[m] = make_uint2(root, __float_as_uint(nodes[2 * (size_t)root + 1].w));
}

void launch_forest_roots(const uint2* range, const float4* nodes, uint32_t n, const uint32_t* toff, uint32_t M,
                         uint2* roots, hipStream_t s) {
    if (M == 0) return;
    k_forest_roots<<<blocks_for(M, 256), 256, 0, s>>>(range, nodes, n, toff, M, roots);
    GI_LAUNCH_CHECK();
}

// --------------------------------------------------------------- traversal


// Vertex i of a mesh as the predicates see it: the moved copy, or for an
// unmoved mesh (no copy written) the source vertices themselves (their
// integer copy when they are on the grid); as integers for grid meshes
// (exact: integer values within 2^26).
__device__ __forceinline__ double4 mesh_vertex(const MeshDev& M, uint32_t i) {
    if (M.v64) return M.v64[i];
    if (M.vi) {  // an unmoved mesh on the grid: its integer copy holds the same values
        const int4 v = M.vi[i];
        return make_double4((double)v.x, (double)v.y, (double)v.z, 0.0);
    }
    const double* p = M.src + 3 * (size_t)i;
    return make_double4(p[0], p[1], p[2], 0.0);
}
__device__ __forceinline__ int4 mesh_vertex_int(const MeshDev& M, uint32_t i) {
    if (M.vi) return M.vi[i];
    const double* p = M.src + 3 * (size_t)i;
    return make_int4((int)p[0], (int)p[1], (int)p[2], 0);
}

__device__ __forceinline__ void flush_stage(uint2* stage, uint32_t fill, uint2* cand, uint32_t cap,
                                            uint32_t* counters) {
    wave_sync_lds();
    uint32_t base = 0;
    if (lane_id() == 0) {
        base = atomicAdd(&counters[CNT_TRAVERSAL_HITS], fill);
        if (u32_add_wraps(base, fill)) atomicOr(&counters[CNT_OVERFLOW], 1u);
    }
    base = __shfl(base, 0);
    for (uint32_t i = lane_id(); i < fill; i += 64) {
        uint32_t pos = base + i;
        if (pos < cap) cand[pos] = stage[i];
    }
    wave_sync_lds();
}

__device__ __forceinline__ bool box_overlap(const float* c, float4 L, float4 H) {
    return (L.x <= c[3]) & (c[0] <= H.x) & (L.y <= c[4]) & (c[1] <= H.y) & (L.z <= c[5]) & (c[2] <= H.z);
}

// Wave-shared descent of k_traverse. When every lane with a walk to do starts
// at the same node, the union box of their queries walks down from it while
// exactly one child of the current node overlaps the union: a child the union
// misses holds no leaf any lane's query overlaps, so every lane's candidates
// lie below the other. The lanes then start their own walks at the node where
// both children overlap the union (a leaf stops the descent too): a lane
// whose own box overlaps both children starts at the left one and ends at the
// node's rope, one that overlaps one child walks that child's subtree only,
// and one that overlaps neither ends there (the two child boxes are already
// loaded). All lanes end at once where the union misses the node or both its
// children. The node index is wave-uniform, so the descent's node loads
// are scalar and leave the vector memory path (whose per-lane return is what
// a lane walk pays per node) to the lanes' own walks. Forest walks take part
// when a wave's queries belong to one mesh (the same start and end rope).
__device__ __forceinline__ void shared_descent(const float4* __restrict__ nodes, uint32_t leafBase, float qlx,
                                               float qly, float qlz, float qhx, float qhy, float qhz,
                                               uint32_t& node, uint32_t& end) {
    const bool walks = node != kSentinel;  // node == end was folded into kSentinel
    if (!__any(walks)) return;
    const uint32_t lead = (uint32_t)__builtin_ctzll(__ballot(walks));
    const uint32_t s0 = __shfl(node, (int)lead), e0 = __shfl(end, (int)lead);
    if (!__all(!walks || (node == s0 && end == e0))) return;
    float c[6] = {walks ? qlx : INFINITY, walks ? qly : INFINITY, walks ? qlz : INFINITY,
                  walks ? qhx : -INFINITY, walks ? qhy : -INFINITY, walks ? qhz : -INFINITY};
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int k = 0; k < 3; ++k) c[k] = fminf(c[k], __shfl_xor(c[k], o));
#pragma unroll
        for (int k = 3; k < 6; ++k) c[k] = fmaxf(c[k], __shfl_xor(c[k], o));
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) c[k] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__float_as_uint(c[k])));
    uint32_t v = __builtin_amdgcn_readfirstlane(s0);
    const uint32_t vend = __builtin_amdgcn_readfirstlane(e0);
    float4 L = nodes[2 * (size_t)v], H = nodes[2 * (size_t)v + 1];
    bool none = !box_overlap(c, L, H);
    uint32_t vrope = vend;
    bool forked = false;
    uint32_t fl = 0, fr = 0;
    float4 flL, flH, frL, frH;
    while (!none && v < leafBase) {
        // children gamma and gamma + 1 (internal or leaf), both loaded at once
        const uint32_t lw = __builtin_amdgcn_readfirstlane(__float_as_uint(L.w));
        const uint32_t left = lw & kNodeIndexMask;
        const uint32_t gamma = left < leafBase ? left : left - leafBase;
        const uint32_t right = (lw & kRightLeaf) ? leafBase + gamma + 1u : gamma + 1u;
        const float4 lL = nodes[2 * (size_t)left], lH = nodes[2 * (size_t)left + 1];
        const float4 rL = nodes[2 * (size_t)right], rH = nodes[2 * (size_t)right + 1];
        const bool ol = box_overlap(c, lL, lH), orr = box_overlap(c, rL, rH);
        if (ol && orr) {
            forked = true;
            fl = left; fr = right;
            flL = lL; flH = lH; frL = rL; frH = rH;
            break;
        }
        if (!ol && !orr) {
            none = true;
            break;
        }
        v = ol ? left : right;
        L = ol ? lL : rL;
        H = ol ? lH : rH;
        vrope = __builtin_amdgcn_readfirstlane(__float_as_uint(H.w));
    }
    if (!walks) return;
    const uint32_t vout = v == s0 ? vend : vrope;  // the rope leaving v
    if (none) {
        node = end = kSentinel;
    } else if (forked) {
        const float q[6] = {qlx, qly, qlz, qhx, qhy, qhz};
        const bool a = box_overlap(q, flL, flH), b = box_overlap(q, frL, frH);
        node = a ? fl : (b ? fr : kSentinel);
        end = a ? (b ? vout : fr) : (b ? vout : kSentinel);  // fr is the left child's rope
    } else {
        node = v;
        end = vout;
    }
}

// A walk handed from a busy lane to an idle one: the rest of a query's walk
// after the subtree its lane is entering, with the query.
struct WalkPiece {
    uint32_t node, end, b, pad;
    float q[6];
};
constexpr uint32_t kPiecesPerWave = 32;  // donors + idle lanes <= 64: at most 32 hand-overs per step

// One thread per triangle of B, in B's Morton order (orderB, so a wave
// queries a compact cluster; B needs no tree of its own), its box read from
// tboxB; stackless traversal of A along the ropes. Overlapping leaf pairs
// are compacted per wave (ballot + popcount) into LDS and flushed to the
// global candidate queue with one atomic per STAGE pairs (the queue chunk).
// Forest (batch): leaf k of B belongs to mesh keysB[k] >> mesh_shift and walks only
// the subtree of the same mesh of A, from its root to the rope leaving it.
//
// Work sharing: a walk is a run of nodes in pre-order, [node, end), and it
// splits at any node v it enters: [v, rope(v)) is v's subtree, [rope(v),
// end) the rest. Most lanes of a wave end after the shared descent or a few
// nodes while a few walk long, so at every step that finds idle lanes, each
// lane entering the children of an overlapping internal node v (up to the
// number of idle lanes) hands [rope(v), end) with its query to an idle lane
// through LDS and keeps v's subtree. The pieces are the entries a stack-based
// traversal would push, spread over the wave; every node of the walk is still
// visited by exactly one lane, so the hits are the same.
template <int TRAV_THREADS, uint32_t TRAV_STAGE>
__global__ __launch_bounds__(TRAV_THREADS) void k_traverse(const float4* __restrict__ nodesA, uint32_t nA,
                                                           const uint32_t* __restrict__ orderB,
                                                           const float4* __restrict__ tboxB, uint32_t nB,
                                                           const uint32_t* __restrict__ keysB, int mesh_shift,
                                                           const uint2* __restrict__ rootsA,
                                                           uint2* __restrict__ cand, uint32_t cap,
                                                           uint32_t* __restrict__ counters) {
    __shared__ uint2 stage_all[TRAV_THREADS / 64][TRAV_STAGE];
    __shared__ WalkPiece pieces_all[TRAV_THREADS / 64][kPiecesPerWave];
    uint2* stage = stage_all[threadIdx.x >> 6];
    WalkPiece* pieces = pieces_all[threadIdx.x >> 6];
    uint32_t k = xcd_block(blockIdx.x, gridDim.x) * TRAV_THREADS + threadIdx.x;
    float qlx = 0.f, qly = 0.f, qlz = 0.f, qhx = 0.f, qhy = 0.f, qhz = 0.f;
    uint32_t b = 0;
    uint32_t node = kSentinel, end = kSentinel;
    if (k < nB) {
        b = orderB[k];
        float4 L = tboxB[2 * (size_t)b], H = tboxB[2 * (size_t)b + 1];
        qlx = L.x; qly = L.y; qlz = L.z;
        qhx = H.x; qhy = H.y; qhz = H.z;
        node = 0;
        if (rootsA) {
            uint2 r = rootsA[keysB[k] >> mesh_shift];
            node = r.x;
            end = r.y;
        }
    }
    const uint32_t leafBase = nA - 1;
    const uint64_t lt = lanemask_lt();
    uint32_t fill = 0;
    if (node == end) node = end = kSentinel;
    shared_descent(nodesA, leafBase, qlx, qly, qlz, qhx, qhy, qhz, node, end);
    // a walk ends at its end rope; the rope chain itself ends in kSentinel
    for (;;) {
        const bool active = node != end && node != kSentinel;
        const uint64_t act = __ballot(active);
        if (act == 0) break;
        bool hit = false, give = false;
        uint32_t a = 0, rest = 0;
        if (active) {
            float4 L = nodesA[2 * (size_t)node], H = nodesA[2 * (size_t)node + 1];
            // all six compares unconditionally (&, not &&): the node's two
            // float4 loads then stay two 16-byte loads
            const bool ov = (L.x <= qhx) & (qlx <= H.x) & (L.y <= qhy) & (qly <= H.y) & (L.z <= qhz) & (qlz <= H.z);
            const uint32_t w0 = __float_as_uint(L.w) & kNodeIndexMask, w1 = __float_as_uint(H.w);
            const bool down = ov && node < leafBase;
            if (ov && node >= leafBase) { hit = true; a = w0; }
            node = down ? w0 : w1;
            give = down && w1 != end && w1 != kSentinel;  // a rest to hand over: [rope, end)
            rest = w1;
        }
        const uint64_t donors = __ballot(give), idle = ~act;
        if (donors != 0 && idle != 0) {
            const uint32_t m = min((uint32_t)__popcll(donors), (uint32_t)__popcll(idle));
            if (give) {
                const uint32_t slot = (uint32_t)__popcll(donors & lt);
                if (slot < m) {
                    WalkPiece& p = pieces[slot];
                    p.node = rest;
                    p.end = end;
                    p.b = b;
                    p.q[0] = qlx; p.q[1] = qly; p.q[2] = qlz;
                    p.q[3] = qhx; p.q[4] = qhy; p.q[5] = qhz;
                    end = rest;  // this lane keeps the subtree it is entering
                }
            }
            wave_sync_lds();
            if (!active) {
                const uint32_t slot = (uint32_t)__popcll(idle & lt);
                if (slot < m) {
                    const WalkPiece& p = pieces[slot];
                    node = p.node;
                    end = p.end;
                    b = p.b;
                    qlx = p.q[0]; qly = p.q[1]; qlz = p.q[2];
                    qhx = p.q[3]; qhy = p.q[4]; qhz = p.q[5];
                }
            }
            wave_sync_lds();
        }
        uint64_t mh = __ballot(hit);
        if (mh) {
            uint32_t cnt = (uint32_t)__popcll(mh);
            if (fill + cnt > TRAV_STAGE) {
                flush_stage(stage, fill, cand, cap, counters);
                fill = 0;
            }
            if (hit) stage[fill + (uint32_t)__popcll(mh & lt)] = make_uint2(a, b);  // a hit lane was active: its query is unchanged
            fill += cnt;
        }
    }
    if (fill) flush_stage(stage, fill, cand, cap, counters);
}

template <int T, uint32_t S>
static void traverse_as(MeshDev A, MeshDev B, const uint32_t* keysB, int mesh_shift, const uint2* rootsA,
                        uint2* cand, uint32_t cap, uint32_t* counters, hipStream_t s) {
    k_traverse<T, S><<<blocks_for(B.nt, T), T, 0, s>>>(A.nodes, A.nt, B.order, B.tbox, B.nt, keysB, mesh_shift, rootsA,
                                                       cand, cap, counters);
}

void launch_traverse(MeshDev A, MeshDev B, const uint32_t* keysB, int mesh_shift, const uint2* rootsA, uint2* cand,
                     uint32_t cap, uint32_t* counters, int variant, hipStream_t s) {
    if (A.nt == 0 || B.nt == 0) return;
    switch (variant) {  // (workgroup size, candidate pairs staged per wave)
        case 0: traverse_as<64, 256>(A, B, keysB, mesh_shift, rootsA, cand, cap, counters, s); break;
        case 1: traverse_as<128, 256>(A, B, keysB, mesh_shift, rootsA, cand, cap, counters, s); break;
        case 3: traverse_as<256, 128>(A, B, keysB, mesh_shift, rootsA, cand, cap, counters, s); break;
        default: traverse_as<256, 256>(A, B, keysB, mesh_shift, rootsA, cand, cap, counters, s); break;
    }
    GI_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ filter


__device__ __forceinline__ uint32_t wave_exclusive_sum(uint32_t v, uint32_t& total) {
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane_id() >= (uint32_t)o) x += y;
    }
    total = __shfl(x, 63);
    return x - v;
}

// Reserves v slots for every thread of a 256-thread workgroup with one atomic
// on `counter` and returns the thread's first slot (exclusive prefix in
// thread order). Same-address atomics from every wave serialise (~10 ns
// each on this GPU), so a stage that appends from all its waves at once
// reserves per workgroup. Every thread of the block must call it.
// A count that wraps past 2^32 sets counters[CNT_OVERFLOW].
__device__ __forceinline__ uint32_t block_reserve(uint32_t v, uint32_t* counters, int which) {
    __shared__ uint32_t wtot[4], wbase[4];
    uint32_t total;
    const uint32_t ex = wave_exclusive_sum(v, total);
    const uint32_t wave = threadIdx.x >> 6;
    if (lane_id() == 0) wtot[wave] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t t = wtot[0] + wtot[1] + wtot[2] + wtot[3];
        uint32_t b = t ? atomicAdd(&counters[which], t) : 0u;
        if (u32_add_wraps(b, t)) atomicOr(&counters[CNT_OVERFLOW], 1u);
        for (int w = 0; w < 4; ++w) { wbase[w] = b; b += wtot[w]; }
    }
    __syncthreads();
    return wbase[wave] + ex;
}

// The same for three counters at once (one pair of barriers): thread slots
// s[j] of v[j] entries each on counters[which[j]].
__device__ __forceinline__ void block_reserve3(const uint32_t (&v)[3], uint32_t* counters, int w0, int w1, int w2,
                                               uint32_t (&slot)[3]) {
    __shared__ uint32_t wtot[3][4], wbase[3][4];
    uint32_t ex[3];
    const uint32_t wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        uint32_t total;
        ex[j] = wave_exclusive_sum(v[j], total);
        if (lane_id() == 0) wtot[j][wave] = total;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int j = (int)threadIdx.x;
        const int which = j == 0 ? w0 : (j == 1 ? w1 : w2);
        const uint32_t t = wtot[j][0] + wtot[j][1] + wtot[j][2] + wtot[j][3];
        uint32_t b = t ? atomicAdd(&counters[which], t) : 0u;
        if (u32_add_wraps(b, t)) atomicOr(&counters[CNT_OVERFLOW], 1u);
        for (int w = 0; w < 4; ++w) { wbase[j][w] = b; b += wtot[j][w]; }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 3; ++j) slot[j] = wbase[j][wave] + ex[j];
}

// Block totals of N values added to counters[which[j]] (one pair of
// barriers; zero totals add nothing).
template <int N>
__device__ __forceinline__ void block_add(const uint32_t (&v)[N], uint32_t* counters, const int (&which)[N]) {
    __shared__ uint32_t red[N][4];
    const uint32_t wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        uint32_t x = v[j];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
        if (lane_id() == 0) red[j][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const uint32_t t = red[j][0] + red[j][1] + red[j][2] + red[j][3];
            if (t) atomicAdd(&counters[which[j]], t);
        }
    }
    __syncthreads();
}

constexpr uint32_t kStatusShift = 30;
constexpr uint32_t kIndexMask = (1u << kStatusShift) - 1u;

__global__ __launch_bounds__(256, 2) void k_filter(MeshDev A, MeshDev B, const uint2* __restrict__ cand, uint32_t cap,
                                                const uint32_t* __restrict__ grid_flags, uint2* __restrict__ exact_list,
                                                uint32_t exact_cap, uint32_t* __restrict__ counters) {
    uint32_t n = counters[CNT_TRAVERSAL_HITS];
    if (n > cap) n = cap;
    const bool grid = grid_flags[0] && grid_flags[1];
    uint32_t c_cand = 0, c_no = 0, c_green = 0, c_yellow = 0;
    int c_pred = 0;
    for (uint32_t base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        uint32_t i = base + threadIdx.x;
        int status = PAIR_NO;
        uint32_t a = 0, b = 0;
        if (i < n) {
            uint2 pr = cand[i];
            a = pr.x;
            b = pr.y;
            const uint32_t* ta = A.tri + 3 * (size_t)a;
            const uint32_t* tb = B.tri + 3 * (size_t)b;
            d3 PA[3], PB[3];
            if (grid) {
                for (int v = 0; v < 3; ++v) {
                    int4 p = mesh_vertex_int(A, ta[v]), q = mesh_vertex_int(B, tb[v]);
                    PA[v] = d3{(double)p.x, (double)p.y, (double)p.z};
                    PB[v] = d3{(double)q.x, (double)q.y, (double)q.z};
                }
            } else {
                for (int v = 0; v < 3; ++v) {
                    double4 p = mesh_vertex(A, ta[v]), q = mesh_vertex(B, tb[v]);
                    PA[v] = d3{p.x, p.y, p.z};
                    PB[v] = d3{q.x, q.y, q.z};
                }
            }
            // exact closed box overlap (the CGAL box_intersection_d candidate set)
            bool ov = fmax(fmax(PA[0].x, PA[1].x), PA[2].x) >= fmin(fmin(PB[0].x, PB[1].x), PB[2].x) &&
                      fmax(fmax(PB[0].x, PB[1].x), PB[2].x) >= fmin(fmin(PA[0].x, PA[1].x), PA[2].x) &&
                      fmax(fmax(PA[0].y, PA[1].y), PA[2].y) >= fmin(fmin(PB[0].y, PB[1].y), PB[2].y) &&
                      fmax(fmax(PB[0].y, PB[1].y), PB[2].y) >= fmin(fmin(PA[0].y, PA[1].y), PA[2].y) &&
                      fmax(fmax(PA[0].z, PA[1].z), PA[2].z) >= fmin(fmin(PB[0].z, PB[1].z), PB[2].z) &&
                      fmax(fmax(PB[0].z, PB[1].z), PB[2].z) >= fmin(fmin(PA[0].z, PA[1].z), PA[2].z);
            if (ov) {
                c_cand++;
                status = classify_pair_fp64(PA, PB, c_pred);
                c_no += status == PAIR_NO;
                c_green += status == PAIR_GREEN;
                c_yellow += status == PAIR_YELLOW;
            }
        }
        const bool emit = status != PAIR_NO;
        const uint32_t slot = block_reserve(emit ? 1u : 0u, counters, CNT_EXACT_LIST);
        if (emit && slot < exact_cap) exact_list[slot] = make_uint2(a, b | ((uint32_t)status << kStatusShift));
    }
    const uint32_t tot[5] = {(uint32_t)c_pred, c_cand, c_no, c_green, c_yellow};
    const int which[5] = {CNT_FILTER_PRED, CNT_CANDIDATES, CNT_FILTER_NO, CNT_FILTER_GREEN, CNT_FILTER_YELLOW};
    block_add<5>(tot, counters, which);
}

void launch_filter(MeshDev A, MeshDev B, const uint2* cand, uint32_t cap, const uint32_t* grid_flags,
                   uint2* exact_list, uint32_t exact_cap, uint32_t* counters, int blocks, hipStream_t s) {
    if (A.nt == 0 || B.nt == 0) return;
    k_filter<<<blocks, 256, 0, s>>>(A, B, cand, cap, grid_flags, exact_list, exact_cap, counters);
    GI_LAUNCH_CHECK();
}

// ------------------------------------------------------------------- exact

typedef ExactTT<GridPolicy> GridExact;

// Per-block tallies of an exact pass, added to the counters once per block.
struct ExactTally {
    uint32_t yes = 0, no = 0, cop = 0, deg = 0, bad = 0, pred = 0;
    __device__ void add(const int status, const int hit, const int coplanar, const int degenerate) {
        if (status == PAIR_YELLOW) { yes += hit; no += !hit; }
        if (status == PAIR_GREEN && !hit) bad++;
        cop += hit && coplanar;
        deg += hit && degenerate;
    }
    __device__ void flush(uint32_t* counters) {
        const uint32_t v[6] = {yes, no, cop, deg, bad, pred};
        const int which[6] = {CNT_EXACT_YELLOW_YES, CNT_EXACT_YELLOW_NO, CNT_COPLANAR, CNT_DEGENERATE, CNT_UNSOUND,
                              CNT_EXACT_PRED};
        block_add<6>(v, counters, which);
    }
};

// Exact decision and segments of the listed pairs of grid meshes. The fast
// instance keeps its per-lane state in registers: a coplanar pair (polygon
// pieces) is only moved to the deferred list, which the full instance then
// processes with the whole result (pieces in per-lane memory). Records and
// segments of both are appended with one atomic per workgroup; a segment
// slot holds 8 int128 (x0 y0 z0 w0 x1 y1 z1 w1).
template <bool kFast>
__global__ __launch_bounds__(256) void k_exact(MeshDev A, MeshDev B, const uint2* __restrict__ list, uint32_t cap,
                                               const uint32_t* __restrict__ grid_flags, int bitsB,
                                               uint64_t* __restrict__ rec_key, uint2* __restrict__ rec_seg,
                                               uint32_t rec_cap, uint64_t* __restrict__ seg_pool_words,
                                               uint32_t seg_cap, uint2* __restrict__ defer, uint32_t defer_cap,
                                               uint32_t* __restrict__ counters) {
    typedef typename std::conditional<kFast, GridExact::Fast, GridExact::Result>::type R;
    if (!(grid_flags[0] && grid_flags[1])) return;  // general doubles: k_exact_wide
    __int128* seg_pool = (__int128*)seg_pool_words;
    uint32_t n = counters[kFast ? CNT_EXACT_LIST : CNT_DEFERRED];
    if (n > cap) n = cap;
    ExactTally tally;
    for (uint32_t base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        uint32_t i = base + threadIdx.x;
        R r;
        r.hit = 0;
        r.nseg = 0;
        r.deferred = 0;
        GridExact::V3 PA[3], PB[3];
        uint32_t a = 0, b = 0;
        uint2 e = make_uint2(0, 0);
        if (i < n) {
            e = list[i];
            a = e.x;
            b = e.y & kIndexMask;
            uint32_t status = e.y >> kStatusShift;
            const uint32_t* ta = A.tri + 3 * (size_t)a;
            const uint32_t* tb = B.tri + 3 * (size_t)b;
            for (int v = 0; v < 3; ++v) {
                int4 p = mesh_vertex_int(A, ta[v]), q = mesh_vertex_int(B, tb[v]);
                PA[v].v[0] = p.x; PA[v].v[1] = p.y; PA[v].v[2] = p.z;
                PB[v].v[0] = q.x; PB[v].v[1] = q.y; PB[v].v[2] = q.z;
            }
            if constexpr (kFast) GridExact::intersect_fast(PA, PB, r);
            else GridExact::intersect(PA, PB, r);
            tally.pred += (uint32_t)r.npred;
            if (!r.deferred) tally.add((int)status, r.hit, r.coplanar, r.degenerate);
        }
        const bool hit = r.hit && !r.deferred;
        const uint32_t nseg = hit ? (uint32_t)r.nseg : 0u;
        const uint32_t want[3] = {r.deferred ? 1u : 0u, hit ? 1u : 0u, nseg};
        uint32_t got[3];
        block_reserve3(want, counters, CNT_DEFERRED, CNT_RESULTS, CNT_SEGMENTS, got);
        const uint32_t slot = got[1], sfirst = got[2];
        if constexpr (kFast) {
            if (r.deferred && got[0] < defer_cap) defer[got[0]] = e;
        }
        if (hit && slot < rec_cap) {
            rec_key[slot] = ((uint64_t)a << bitsB) | (uint64_t)b;
            rec_seg[slot] = make_uint2(sfirst, nseg);
            for (uint32_t sI = 0; sI < nseg; ++sI) {
                uint32_t dst = sfirst + sI;
                if (dst >= seg_cap) break;
                GridExact::HP p, q;
                if constexpr (kFast) {
                    p = r.p;
                    q = r.q;
                } else {
                    GridExact::segment(PA, PB, r, (int)sI, p, q);
                }
                __int128* o = seg_pool + 8 * (size_t)dst;
                o[0] = p.h[0]; o[1] = p.h[1]; o[2] = p.h[2]; o[3] = p.w;
                o[4] = q.h[0]; o[5] = q.h[1]; o[6] = q.h[2]; o[7] = q.w;
            }
        }
    }
    tally.flush(counters);
}

// ----------------------------------------------- grid pairs, one lane each
//
// Every pair of grid meshes but the coplanar ones and those with a
// degenerate triangle. T_A ^ T_B is then the overlap, on the line L =
// plane_A ^ plane_B, of the segments where each triangle meets the other's
// plane (tri_plane_segment below), whose end points are exact homogeneous
// int128 points. On an axis k where L's direction nA x nB is non-zero,
// points of L are ordered by their k coordinate alone: the overlap is [max of
// the two lower ends, min of the two upper ends], empty when they cross.
// Each comparison of h_k / w values first runs in FP64 with a certified bound
// and falls back to the exact cross-multiplied one (256-bit products) when
// the bound does not decide. One lane per pair: the 8-lane kernel below
// computes the shared part of a pair in all of its lanes. The coplanar and
// degenerate pairs go to the slow list, which that kernel then takes.

// Exact numbers to double: |error| <= 4 eps |x| for int128 (the magnitude's
// two 64-bit halves each rounded once, then their sum) and <= 8 eps |x| for
// the 256-bit numbers (four limbs, three sums of non-negative terms), eps =
// 2^-53.
__device__ __forceinline__ double exact_to_f64(i128 x) {
    const u128 u = uabs128(x);
    const double d = (double)(uint64_t)(u >> 64) * 18446744073709551616.0 + (double)(uint64_t)u;
    return x < 0 ? -d : d;
}
__device__ __forceinline__ double exact_to_f64(const W256& x) {
    const W256 u = w256_abs(x);
    const double hi = (double)u.w[3] * 6.277101735386681e57 + (double)u.w[2] * 3.402823669209385e38;  // 2^192, 2^128
    const double lo = (double)u.w[1] * 18446744073709551616.0 + (double)u.w[0];
    const double d = hi + lo;
    return w256_neg(x) ? -d : d;
}

// Sign of xa - xb where each x is the FP64 quotient h / w of exact values
// converted as above (so within 2 * 8 + 1 = 17 eps of h / w, relative), or 0
// when that error could flip it: the bound is 2^-46 = 128 eps of the
// magnitudes, well above the 34 eps the two quotients may be off together.
__device__ __forceinline__ int cmp_f64_certified(double xa, double xb) {
    const double d = xa - xb;
    const double bound = 1.4210854715202004e-14 * (fabs(xa) + fabs(xb));  // 2^-46
    return d > bound ? 1 : (d < -bound ? -1 : 0);
}

// The segment where triangle (T0, T1, T2) meets the other triangle's plane,
// from its vertices' orientations o0..o2 against that plane (signs s0..s2,
// neither all zero nor all of one sign): with no zero sign, the points where
// the two edges at the vertex alone on its side cross the plane; with one
// zero, that vertex and, when the other two lie on opposite sides, the point
// where their edge crosses (else the vertex alone, p = q); with two zeros,
// the edge between them. Vertices and values come as separate arguments and
// are picked by selects: a select between elements of one array would be
// folded into an indexed access of a per-lane array in scratch memory.
template <class E>
__device__ __forceinline__ void tri_plane_segment(typename E::V3 T0, typename E::V3 T1, typename E::V3 T2,
                                                  typename E::B o0, typename E::B o1, typename E::B o2, int s0, int s1,
                                                  int s2, typename E::HP& p, typename E::HP& q) {
    typedef typename E::C C;
    typedef typename E::B B;
    const int zeros = (s0 == 0) + (s1 == 0) + (s2 == 0);
    // the vertex U the segment hangs on: the apex (no zero), the zero vertex
    // (one zero) or the first zero vertex (two zeros); V, W the others in
    // cyclic order
    const int ap = zeros == 0 ? (s0 == s1 ? 2 : (s0 == s2 ? 1 : 0))
                              : (s0 == 0 ? (zeros == 2 && s1 != 0 ? 2 : 0) : (s1 == 0 ? 1 : 2));
    typename E::V3 U, V, W;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const C t0 = T0.v[c], t1 = T1.v[c], t2 = T2.v[c];
        U.v[c] = ap == 0 ? t0 : (ap == 1 ? t1 : t2);
        V.v[c] = ap == 0 ? t1 : (ap == 1 ? t2 : t0);
        W.v[c] = ap == 0 ? t2 : (ap == 1 ? t0 : t1);
    }
    const B ou = ap == 0 ? o0 : (ap == 1 ? o1 : o2);
    const B ov = ap == 0 ? o1 : (ap == 1 ? o2 : o0);
    const B ow = ap == 0 ? o2 : (ap == 1 ? o0 : o1);
    const int sv = ap == 0 ? s1 : (ap == 1 ? s2 : s0), sw = ap == 0 ? s2 : (ap == 1 ? s0 : s1);
    if (zeros == 0) {
        p = E::hpCrossing(U, V, ou, ov);
        q = E::hpCrossing(U, W, ou, ow);
    } else if (zeros == 1) {
        p = E::hpInt(U);
        q = sv == sw ? p : E::hpCrossing(V, W, ov, ow);
    } else {  // U and the other zero vertex (V or W) span the segment
        p = E::hpInt(U);
        q = E::hpInt(sv == 0 ? V : W);
    }
}

// A pair of non-degenerate, non-coplanar triangles: T_A ^ T_B is the
// overlap, on the line L = plane_A ^ plane_B, of the segments where each
// triangle meets the other's plane. Returns -1 when a triangle's vertices
// all lie on the other's plane (coplanar pair, or a degenerate triangle: a
// zero normal puts the other triangle's vertices all on its "plane"), left
// to the general passes; 0 when the triangles do not meet; 1 with the
// segment [lo, hi] (ordered along an axis of L; lo = hi for a point). When
// neither triangle has all its vertices on the other's plane, the planes
// are not parallel (a parallel plane through a vertex of the other triangle
// would be that triangle's plane) or the triangles lie apart, so L exists.
// npred counts the signs evaluated.
template <class E, class P>
__device__ __forceinline__ int transversal_segment(const typename E::V3 (&PA)[3], const typename E::V3 (&PB)[3],
                                                   typename E::HP& lo, typename E::HP& hi, uint32_t& npred) {
    typedef typename E::V3 V3;
    typedef typename E::HP HP;
    typedef typename E::B B;
    const V3 nA = E::cross(E::sub(PA[1], PA[0]), E::sub(PA[2], PA[0]));
    const V3 nB = E::cross(E::sub(PB[1], PB[0]), E::sub(PB[2], PB[0]));
    B oA[3], oB[3];
    int sA[3], sB[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        oB[v] = E::dotB(nA, E::sub(PB[v], PA[0]));  // B's vertices against plane A
        sB[v] = P::sgnB(oB[v]);
        oA[v] = E::dotB(nB, E::sub(PA[v], PB[0]));  // A's against plane B
        sA[v] = P::sgnB(oA[v]);
    }
    npred += 6;
    if ((sA[0] == 0 && sA[1] == 0 && sA[2] == 0) || (sB[0] == 0 && sB[1] == 0 && sB[2] == 0)) return -1;
    if ((sA[0] == sA[1] && sA[1] == sA[2]) || (sB[0] == sB[1] && sB[1] == sB[2])) return 0;  // all on one side
    HP pa0, pa1, pb0, pb1;
    tri_plane_segment<E>(PA[0], PA[1], PA[2], oA[0], oA[1], oA[2], sA[0], sA[1], sA[2], pa0, pa1);
    tri_plane_segment<E>(PB[0], PB[1], PB[2], oB[0], oB[1], oB[2], sB[0], sB[1], sB[2], pb0, pb1);
    // axis of the largest component of L's direction nA x nB
    const B d0 = P::mulCC(nA.v[1], nB.v[2]) - P::mulCC(nA.v[2], nB.v[1]);
    const B d1 = P::mulCC(nA.v[2], nB.v[0]) - P::mulCC(nA.v[0], nB.v[2]);
    const B d2 = P::mulCC(nA.v[0], nB.v[1]) - P::mulCC(nA.v[1], nB.v[0]);
    int k = P::cmpAbsB(d1, d0) > 0 ? 1 : 0;
    if (P::cmpAbsB(d2, k ? d1 : d0) > 0) k = 2;
    auto xk = [&](const HP& Q) { return exact_to_f64(E::hcomp(Q, k)) / exact_to_f64(Q.w); };
    double xa0 = xk(pa0), xa1 = xk(pa1), xb0 = xk(pb0), xb1 = xk(pb1);
    auto cmp = [&](const HP& Q, double xq, const HP& R, double xr) {
        int c = cmp_f64_certified(xq, xr);
        if (c == 0) c = E::hpCmpAxis(Q, R, k);
        return c;
    };
    npred += 5;
    if (cmp(pa0, xa0, pa1, xa1) > 0) {  // each triangle's pair in order along k
        const HP t = pa0; pa0 = pa1; pa1 = t;
        const double x = xa0; xa0 = xa1; xa1 = x;
    }
    if (cmp(pb0, xb0, pb1, xb1) > 0) {
        const HP t = pb0; pb0 = pb1; pb1 = t;
        const double x = xb0; xb0 = xb1; xb1 = x;
    }
    const bool la = cmp(pa0, xa0, pb0, xb0) >= 0, ua = cmp(pa1, xa1, pb1, xb1) <= 0;
    lo = la ? pa0 : pb0;
    hi = ua ? pa1 : pb1;
    return cmp(lo, la ? xa0 : xb0, hi, ua ? xa1 : xb1) <= 0 ? 1 : 0;
}

// Lane (lane ^ 1)'s copy of a value through DPP quad permutes (no LDS
// round trip): the two lanes of a pair trade what each computed. Both lanes
// of every pair must be active at the call.
__device__ __forceinline__ int pair_swap(int x) {
    return __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);  // quad_perm [1, 0, 3, 2]
}
__device__ __forceinline__ uint64_t pair_swap(uint64_t x) {
    const uint32_t lo = (uint32_t)pair_swap((int)(uint32_t)x), hi = (uint32_t)pair_swap((int)(uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ int64_t pair_swap(int64_t x) { return (int64_t)pair_swap((uint64_t)x); }
__device__ __forceinline__ double pair_swap(double x) {
    return __builtin_bit_cast(double, pair_swap(__builtin_bit_cast(uint64_t, x)));
}
__device__ __forceinline__ i128 pair_swap(i128 x) {
    const u128 u = (u128)x;
    return (i128)(((u128)pair_swap((uint64_t)(u >> 64)) << 64) | (u128)pair_swap((uint64_t)u));
}
__device__ __forceinline__ W256 pair_swap(const W256& x) {
    W256 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r.w[i] = pair_swap(x.w[i]);
    return r;
}
template <class E>
__device__ __forceinline__ typename E::V3 pair_swap_v3(const typename E::V3& x) {
    typename E::V3 r;
#pragma unroll
    for (int c = 0; c < 3; ++c) r.v[c] = pair_swap(x.v[c]);
    return r;
}
template <class E>
__device__ __forceinline__ typename E::HP pair_swap_hp(const typename E::HP& x) {
    typename E::HP r;
#pragma unroll
    for (int c = 0; c < 3; ++c) r.h[c] = pair_swap(x.h[c]);
    r.w = pair_swap(x.w);
    return r;
}

// transversal_segment on two lanes: the even lane of the pair holds
// triangle A (T), the odd lane triangle B. Each lane orients its own
// vertices against the other triangle's plane (normal and first vertex
// traded), builds its own triangle's segment on that plane and puts it in
// order along L's axis (nA x nB and nB x nA give the same axis); the odd
// lane's segment then goes to the even lane, whose return value, lo and hi
// are those of transversal_segment for (A, B). The return value is -1 or 0
// on both lanes alike when the early tests decide the pair; otherwise only
// the even lane's result holds. npred as there.
template <class E, class P>
__device__ __forceinline__ int transversal_segment_pair(const typename E::V3 (&T)[3], typename E::HP& lo,
                                                        typename E::HP& hi, uint32_t& npred) {
    typedef typename E::V3 V3;
    typedef typename E::HP HP;
    typedef typename E::B B;
    const V3 n = E::cross(E::sub(T[1], T[0]), E::sub(T[2], T[0]));
    const V3 nO = pair_swap_v3<E>(n);
    const V3 O0 = pair_swap_v3<E>(T[0]);
    B o[3];
    int s[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        o[v] = E::dotB(nO, E::sub(T[v], O0));  // own vertices against the other plane
        s[v] = P::sgnB(o[v]);
    }
    npred += 6;
    const int code = (s[0] == 0 && s[1] == 0 && s[2] == 0) ? 2 : ((s[0] == s[1] && s[1] == s[2]) ? 1 : 0);
    const int codeO = pair_swap(code);
    if (code == 2 || codeO == 2) return -1;
    if (code == 1 || codeO == 1) return 0;  // all on one side
    HP p, q;
    tri_plane_segment<E>(T[0], T[1], T[2], o[0], o[1], o[2], s[0], s[1], s[2], p, q);
    // axis of the largest component of L's direction (n x nO = +-nA x nB)
    const B d0 = P::mulCC(n.v[1], nO.v[2]) - P::mulCC(n.v[2], nO.v[1]);
    const B d1 = P::mulCC(n.v[2], nO.v[0]) - P::mulCC(n.v[0], nO.v[2]);
    const B d2 = P::mulCC(n.v[0], nO.v[1]) - P::mulCC(n.v[1], nO.v[0]);
    int k = P::cmpAbsB(d1, d0) > 0 ? 1 : 0;
    if (P::cmpAbsB(d2, k ? d1 : d0) > 0) k = 2;
    auto xk = [&](const HP& Q) { return exact_to_f64(E::hcomp(Q, k)) / exact_to_f64(Q.w); };
    auto cmp = [&](const HP& Q, double xq, const HP& R, double xr) {
        int c = cmp_f64_certified(xq, xr);
        if (c == 0) c = E::hpCmpAxis(Q, R, k);
        return c;
    };
    npred += 5;
    double xp = xk(p), xq = xk(q);
    if (cmp(p, xp, q, xq) > 0) {  // own pair in order along k
        const HP t = p; p = q; q = t;
        const double x = xp; xp = xq; xq = x;
    }
    // the other lane's ordered pair (on the even lane: B's)
    const HP pO = pair_swap_hp<E>(p), qO = pair_swap_hp<E>(q);
    const double xpO = pair_swap(xp), xqO = pair_swap(xq);
    const bool la = cmp(p, xp, pO, xpO) >= 0, ua = cmp(q, xq, qO, xqO) <= 0;
    lo = la ? p : pO;
    hi = ua ? q : qO;
    return cmp(lo, la ? xp : xpO, hi, ua ? xq : xqO) <= 0 ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_exact_grid_fast(MeshDev A, MeshDev B, const uint2* __restrict__ list,
                                                         uint32_t cap, const uint32_t* __restrict__ grid_flags,
                                                         int bitsB, uint64_t* __restrict__ rec_key,
                                                         uint2* __restrict__ rec_seg, uint32_t rec_cap,
                                                         uint64_t* __restrict__ seg_pool_words, uint32_t seg_cap,
                                                         uint2* __restrict__ slow, uint32_t slow_cap,
                                                         uint32_t* __restrict__ counters) {
    typedef GridExact::V3 V3;
    typedef GridExact::HP HP;
    if (!(grid_flags[0] && grid_flags[1])) return;  // general doubles: k_exact_wide_fast
    __int128* seg_pool = (__int128*)seg_pool_words;
    uint32_t n = counters[CNT_EXACT_LIST];
    if (n > cap) n = cap;
    ExactTally tally;
    // two lanes per pair: the even lane holds triangle A and writes the results
    const bool sideA = (threadIdx.x & 1u) == 0u;
    for (uint32_t base = blockIdx.x * 128; base < n; base += gridDim.x * 128) {
        const uint32_t i = base + (threadIdx.x >> 1);
        bool to_slow = false, hit = false;
        uint2 en = make_uint2(0, 0);
        HP lo, hi;
        if (i < n) {
            en = list[i];
            const uint32_t* t = sideA ? A.tri + 3 * (size_t)en.x : B.tri + 3 * (size_t)(en.y & kIndexMask);
            const int4* vi = sideA ? A.vi : B.vi;
            const double* src = sideA ? A.src : B.src;
            V3 T[3];
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                const uint32_t j = t[v];
                if (vi) {
                    const int4 p = vi[j];
                    T[v].v[0] = p.x; T[v].v[1] = p.y; T[v].v[2] = p.z;
                } else {
                    const double* p = src + 3 * (size_t)j;
                    T[v].v[0] = (int64_t)p[0]; T[v].v[1] = (int64_t)p[1]; T[v].v[2] = (int64_t)p[2];
                }
            }
            uint32_t np = 0;
            const int r = transversal_segment_pair<GridExact, GridPolicy>(T, lo, hi, np);
            to_slow = sideA && r < 0;
            hit = sideA && r > 0;
            if (sideA) {
                tally.pred += np;
                if (!to_slow) tally.add((int)(en.y >> kStatusShift), hit, 0, 0);
            }
        }
        const uint32_t want[3] = {to_slow ? 1u : 0u, hit ? 1u : 0u, hit ? 1u : 0u};
        uint32_t got[3];
        block_reserve3(want, counters, CNT_SLOW, CNT_RESULTS, CNT_SEGMENTS, got);
        if (to_slow && got[0] < slow_cap) slow[got[0]] = en;
        if (hit && got[1] < rec_cap) {
            rec_key[got[1]] = ((uint64_t)en.x << bitsB) | (uint64_t)(en.y & kIndexMask);
            rec_seg[got[1]] = make_uint2(got[2], 1u);
            if (got[2] < seg_cap) {
                __int128* o = seg_pool + 8 * (size_t)got[2];
                o[0] = lo.h[0]; o[1] = lo.h[1]; o[2] = lo.h[2]; o[3] = lo.w;
                o[4] = hi.h[0]; o[5] = hi.h[1]; o[6] = hi.h[2]; o[7] = hi.w;
            }
        }
    }
    tally.flush(counters);
}

// ----------------------------------------------- grid pairs, 8 lanes each
//
// The transversal pairs of grid meshes (the common case: neither triangle
// degenerate, not coplanar) decided and constructed by 8 lanes per pair, 8
// pairs per wave: every lane of a pair's group computes the two normals, the
// six plane orientations and the axis of the line L = plane_A ^ plane_B; lane
// e < 6 then takes edge e (A's edges 0-2 against triangle B, B's against A)
// with the template's edgePiece, and three butterfly steps over the group
// keep the extreme points along that axis, which are the segment's ends
// (ExactTT::intersect_fast does the same edges one after another in one
// lane). Degenerate and coplanar pairs, and pairs with an edge lying in the
// other triangle's plane, go to the deferred list for the full pass. The
// group's first lane writes the record and the segment.
constexpr uint32_t kPairLanes = 8;

__device__ __forceinline__ i128 shfl_xor_i128(i128 x, int m) {
    const u128 u = (u128)x;
    const unsigned long long lo = __shfl_xor((unsigned long long)(uint64_t)u, m);
    const unsigned long long hi = __shfl_xor((unsigned long long)(uint64_t)(u >> 64), m);
    return (i128)(((u128)hi << 64) | (u128)lo);
}

__device__ __forceinline__ GridExact::HP shfl_xor_hp(const GridExact::HP& p, int m) {
    GridExact::HP r;
#pragma unroll
    for (int k = 0; k < 3; ++k) r.h[k] = shfl_xor_i128(p.h[k], m);
    r.w = shfl_xor_i128(p.w, m);
    return r;
}

__global__ __launch_bounds__(256) void k_exact_grid_lanes(MeshDev A, MeshDev B, const uint2* __restrict__ list,
                                                          uint32_t cap, const uint32_t* __restrict__ grid_flags,
                                                          int bitsB, uint64_t* __restrict__ rec_key,
                                                          uint2* __restrict__ rec_seg, uint32_t rec_cap,
                                                          uint64_t* __restrict__ seg_pool_words, uint32_t seg_cap,
                                                          uint2* __restrict__ defer, uint32_t defer_cap,
                                                          uint32_t* __restrict__ counters, int list_counter) {
    typedef GridExact::V3 V3;
    typedef GridExact::HP HP;
    typedef GridExact::B Bn;
    if (!(grid_flags[0] && grid_flags[1])) return;  // general doubles: k_exact_wide
    __int128* seg_pool = (__int128*)seg_pool_words;
    uint32_t n = counters[list_counter];
    if (n > cap) n = cap;
    constexpr uint32_t kPerBlock = 256 / kPairLanes;
    const int e = (int)(lane_id() & (kPairLanes - 1));
    const bool leader = e == 0;
    ExactTally tally;
    for (uint32_t base = blockIdx.x * kPerBlock; base < n; base += gridDim.x * kPerBlock) {
        const uint32_t i = base + threadIdx.x / kPairLanes;
        const bool valid = i < n;
        uint2 en = make_uint2(0, 0);
        uint32_t a = 0, b = 0, status = 0;
        V3 PA[3], PB[3];
        if (valid) {
            en = list[i];
            a = en.x;
            b = en.y & kIndexMask;
            status = en.y >> kStatusShift;
            const uint32_t* ta = A.tri + 3 * (size_t)a;
            const uint32_t* tb = B.tri + 3 * (size_t)b;
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                const int4 p = mesh_vertex_int(A, ta[v]), q = mesh_vertex_int(B, tb[v]);
                PA[v].v[0] = p.x; PA[v].v[1] = p.y; PA[v].v[2] = p.z;
                PB[v].v[0] = q.x; PB[v].v[1] = q.y; PB[v].v[2] = q.z;
            }
        } else {
#pragma unroll
            for (int v = 0; v < 3; ++v)
#pragma unroll
                for (int k = 0; k < 3; ++k) PA[v].v[k] = PB[v].v[k] = 0;
        }
        // shared by the group: normals, shapes, plane orientations, axis
        const V3 nA = GridExact::cross(GridExact::sub(PA[1], PA[0]), GridExact::sub(PA[2], PA[0]));
        const V3 nB = GridExact::cross(GridExact::sub(PB[1], PB[0]), GridExact::sub(PB[2], PB[0]));
        bool live = valid, deferred = false;
        int shared_pred = 0;
        if (live && (GridExact::isZero(nA) || GridExact::isZero(nB))) {  // a degenerate triangle
            deferred = true;
            live = false;
        }
        Bn oA[3], oB[3];
        int sA[3], sB[3];
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            oB[v] = GridExact::dotB(nA, GridExact::sub(PB[v], PA[0]));
            sB[v] = sgn128(oB[v]);
            oA[v] = GridExact::dotB(nB, GridExact::sub(PA[v], PB[0]));
            sA[v] = sgn128(oA[v]);
        }
        if (live) {
            shared_pred += 3;
            if (sB[0] == 0 && sB[1] == 0 && sB[2] == 0) {  // coplanar
                deferred = true;
                live = false;
            } else if (sB[0] * sB[1] > 0 && sB[1] * sB[2] > 0) {
                live = false;
            } else {
                shared_pred += 3;
                if (sA[0] * sA[1] > 0 && sA[1] * sA[2] > 0) live = false;
            }
        }
        const Bn d0 = (Bn)nA.v[1] * nB.v[2] - (Bn)nA.v[2] * nB.v[1];
        const Bn d1 = (Bn)nA.v[2] * nB.v[0] - (Bn)nA.v[0] * nB.v[2];
        const Bn d2 = (Bn)nA.v[0] * nB.v[1] - (Bn)nA.v[1] * nB.v[0];
        int k = GridPolicy::cmpAbsB(d1, d0) > 0 ? 1 : 0;
        if (GridPolicy::cmpAbsB(d2, k ? d1 : d0) > 0) k = 2;
        // this lane's edge
        HP lo, hi;
        int cnt = 0, np = 0, in_plane = 0;
        if (live && e < 6) {
            // the lane's edge (Ti, Tj), the other triangle O, by component
            // selects (arrays picked at run time would live in scratch memory)
            const bool eA = e < 3;
            V3 Ti, Tj, O[3], nO;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int64_t a0 = PA[0].v[c], a1 = PA[1].v[c], a2 = PA[2].v[c];
                const int64_t b0 = PB[0].v[c], b1 = PB[1].v[c], b2 = PB[2].v[c];
                Ti.v[c] = e == 0 ? a0 : e == 1 ? a1 : e == 2 ? a2 : e == 3 ? b0 : e == 4 ? b1 : b2;
                Tj.v[c] = e == 0 ? a1 : e == 1 ? a2 : e == 2 ? a0 : e == 3 ? b1 : e == 4 ? b2 : b0;
                O[0].v[c] = eA ? b0 : a0;
                O[1].v[c] = eA ? b1 : a1;
                O[2].v[c] = eA ? b2 : a2;
                nO.v[c] = eA ? nB.v[c] : nA.v[c];
            }
            const Bn oi = e == 0 ? oA[0] : e == 1 ? oA[1] : e == 2 ? oA[2] : e == 3 ? oB[0] : e == 4 ? oB[1] : oB[2];
            const Bn oj = e == 0 ? oA[1] : e == 1 ? oA[2] : e == 2 ? oA[0] : e == 3 ? oB[1] : e == 4 ? oB[2] : oB[0];
            const int si = sgn128(oi), sj = sgn128(oj);
            if (si == 0 && sj == 0) {
                in_plane = 1;  // an edge in the other triangle's plane: the full pass clips it
            } else {
                HP p0, p1;
                const int m = GridExact::edgePiece(Ti, Tj, O, nO, oi, oj, si, sj, p0, p1, np);
                if (m >= 1) {
                    lo = p0;
                    hi = p0;
                    cnt = 1;
                }
            }
        }
        // a pair with an edge in the other's plane goes to the full pass whole
#pragma unroll
        for (int off = 1; off < (int)kPairLanes; off <<= 1) in_plane |= __shfl_xor(in_plane, off);
        if (in_plane) {
            deferred = true;
            live = false;
            cnt = 0;
        }
        // extremes over the group (every lane takes part in the shuffles)
#pragma unroll
        for (int off = 1; off < (int)kPairLanes; off <<= 1) {
            const int ocnt = __shfl_xor(cnt, off);
            const HP olo = shfl_xor_hp(lo, off), ohi = shfl_xor_hp(hi, off);
            if (ocnt) {
                if (!cnt) {
                    lo = olo;
                    hi = ohi;
                } else {
                    if (GridExact::hpCmpAxis(olo, lo, k) < 0) lo = olo;
                    if (GridExact::hpCmpAxis(ohi, hi, k) > 0) hi = ohi;
                }
                cnt += ocnt;
            }
        }
        const bool hit = live && cnt > 0;
        tally.pred += (uint32_t)(np + (leader ? shared_pred : 0));
        if (leader && valid && !deferred) tally.add((int)status, hit, 0, 0);
        const bool dput = leader && deferred;
        const bool put = leader && hit;
        const uint32_t want[3] = {dput ? 1u : 0u, put ? 1u : 0u, put ? 1u : 0u};
        uint32_t got[3];
        block_reserve3(want, counters, CNT_DEFERRED, CNT_RESULTS, CNT_SEGMENTS, got);
        const uint32_t dslot = got[0], slot = got[1], sfirst = got[2];
        if (dput && dslot < defer_cap) defer[dslot] = en;
        if (put && slot < rec_cap) {
            rec_key[slot] = ((uint64_t)a << bitsB) | (uint64_t)b;
            rec_seg[slot] = make_uint2(sfirst, 1u);
            if (sfirst < seg_cap) {
                __int128* o = seg_pool + 8 * (size_t)sfirst;
                o[0] = lo.h[0]; o[1] = lo.h[1]; o[2] = lo.h[2]; o[3] = lo.w;
                o[4] = hi.h[0]; o[5] = hi.h[1]; o[6] = hi.h[2]; o[7] = hi.w;
            }
        }
    }
    tally.flush(counters);
}

__device__ __forceinline__ void store_w256(uint64_t* __restrict__ o, const W256& x) {
    o[0] = x.w[0]; o[1] = x.w[1]; o[2] = x.w[2]; o[3] = x.w[3];
}

// The same for general doubles: each pair is moved into its own integer
// frame (exact_wide.h) and decided and constructed there on int128/256-bit
// numbers; a segment slot holds 8 numbers of 4 limbs and the frame's
// power-of-two scale (seg_exp). A pair whose frame or endpoints do not fit
// goes to the residual list for the host stage, and counts nowhere else.
template <bool kFast>
__global__ __launch_bounds__(256) void k_exact_wide(MeshDev A, MeshDev B, const uint2* __restrict__ list,
                                                    uint32_t cap, const uint32_t* __restrict__ grid_flags, int bitsB,
                                                    uint64_t* __restrict__ rec_key, uint2* __restrict__ rec_seg,
                                                    uint32_t rec_cap, uint64_t* __restrict__ seg_pool,
                                                    int32_t* __restrict__ seg_exp, uint32_t seg_cap,
                                                    uint2* __restrict__ defer, uint32_t defer_cap,
                                                    uint2* __restrict__ resid, uint32_t resid_cap,
                                                    uint32_t* __restrict__ counters, int list_counter) {
    typedef typename std::conditional<kFast, WideExact::Fast, WideExact::Result>::type R;
    if (grid_flags[0] && grid_flags[1]) return;  // grid meshes: k_exact_grid_fast
    uint32_t n = counters[list_counter];
    if (n > cap) n = cap;
    ExactTally tally;
    for (uint32_t base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        uint32_t i = base + threadIdx.x;
        R r;
        r.hit = 0;
        r.nseg = 0;
        r.deferred = 0;
        WideExact::V3 PA[3], PB[3];
        WideFrame fr;
        fr.exp2 = 0;
        fr.shifted = 0;
        bool residual = false;
        uint32_t a = 0, b = 0;
        uint2 e = make_uint2(0, 0);
        if (i < n) {
            e = list[i];
            a = e.x;
            b = e.y & kIndexMask;
            const uint32_t status = e.y >> kStatusShift;
            const uint32_t* ta = A.tri + 3 * (size_t)a;
            const uint32_t* tb = B.tri + 3 * (size_t)b;
            double pa[9], pb[9];
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                const double4 p = mesh_vertex(A, ta[v]), q = mesh_vertex(B, tb[v]);
                pa[3 * v] = p.x; pa[3 * v + 1] = p.y; pa[3 * v + 2] = p.z;
                pb[3 * v] = q.x; pb[3 * v + 1] = q.y; pb[3 * v + 2] = q.z;
            }
            residual = !wide_frame(pa, pb, PA, PB, fr);
            if (!residual) {
                if constexpr (kFast) {
                    WideExact::intersect_fast(PA, PB, r);
                    if (r.hit && !r.deferred) residual = !wide_absolute(fr, r.p) || !wide_absolute(fr, r.q);
                } else {
                    WideExact::intersect(PA, PB, r);
                    for (int sI = 0; sI < r.nseg && !residual; ++sI) {
                        WideExact::HP p, q;
                        WideExact::segment(PA, PB, r, sI, p, q);
                        residual = !wide_absolute(fr, p) || !wide_absolute(fr, q);
                    }
                }
            }
            if (residual) {
                r.hit = 0;
                r.deferred = 0;
            } else {
                tally.pred += (uint32_t)r.npred;
                if (!r.deferred) tally.add((int)status, r.hit, r.coplanar, r.degenerate);
            }
        }
        const uint32_t rslot = block_reserve(residual ? 1u : 0u, counters, CNT_RESIDUAL);
        if (residual && rslot < resid_cap) resid[rslot] = e;
        if constexpr (kFast) {
            const uint32_t dslot = block_reserve(r.deferred ? 1u : 0u, counters, CNT_DEFERRED);
            if (r.deferred && dslot < defer_cap) defer[dslot] = e;
        }
        const bool hit = r.hit && !r.deferred;
        const uint32_t slot = block_reserve(hit ? 1u : 0u, counters, CNT_RESULTS);
        const uint32_t nseg = hit ? (uint32_t)r.nseg : 0u;
        const uint32_t sfirst = block_reserve(nseg, counters, CNT_SEGMENTS);
        if (hit && slot < rec_cap) {
            rec_key[slot] = ((uint64_t)a << bitsB) | (uint64_t)b;
            rec_seg[slot] = make_uint2(sfirst, nseg);
            for (uint32_t sI = 0; sI < nseg; ++sI) {
                const uint32_t dst = sfirst + sI;
                if (dst >= seg_cap) break;
                WideExact::HP p, q;
                if constexpr (kFast) {
                    p = r.p;
                    q = r.q;
                } else {
                    WideExact::segment(PA, PB, r, (int)sI, p, q);
                    wide_absolute(fr, p);
                    wide_absolute(fr, q);
                }
                uint64_t* o = seg_pool + (size_t)8 * kSegLimbsWide * dst;
                store_w256(o + 0, p.h[0]); store_w256(o + 4, p.h[1]); store_w256(o + 8, p.h[2]);
                store_w256(o + 12, p.w);
                store_w256(o + 16, q.h[0]); store_w256(o + 20, q.h[1]); store_w256(o + 24, q.h[2]);
                store_w256(o + 28, q.w);
                seg_exp[dst] = fr.exp2;
            }
        }
    }
    tally.flush(counters);
}

// The pairs of general doubles, one lane each: the pair in its integer
// frame (exact_wide.h), transversal_segment on 256-bit numbers, the end
// points moved back out of the frame. Pairs whose frame or end points do not
// fit go to the host residual list; coplanar and degenerate pairs to the slow
// list (k_exact_wide<true> on it).
// wide_frame on the two lanes of a pair, each with its own triangle's 9
// doubles (t, vertex-major; the even lane triangle A): the lanes trade the
// lowest exponent, the widest coordinate, A's vertex 0 (the translation)
// and the widest difference, so both come to the frame wide_frame gives for
// all 18 coordinates, and T holds the lane's own triangle in it.
__device__ __forceinline__ bool wide_frame_pair(const double (&t)[9], bool sideA, WideExact::V3 (&T)[3],
                                                WideFrame& f) {
    int64_t m[9];
    int q[9];
    int ok = 1, emin = 0x7FFFFFFF;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        if (!dyadic_parts(t[i], m[i], q[i])) {
            ok = 0;
            m[i] = 0;
            q[i] = 0;
        }
        if (m[i] != 0 && q[i] < emin) emin = q[i];
    }
    emin = min(emin, pair_swap(emin));
    if (emin == 0x7FFFFFFF) emin = 0;  // all zero
    i128 X[9];
    int kmax = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        X[i] = 0;
        if (m[i] == 0) continue;
        const uint64_t um = m[i] < 0 ? (uint64_t)0 - (uint64_t)m[i] : (uint64_t)m[i];
        const int top = 63 - __builtin_clzll(um) + q[i] - emin;  // position of the top bit of |X|
        if (top > 125) {
            ok = 0;
            continue;
        }
        X[i] = (i128)m[i] * ((i128)1 << (q[i] - emin));
        kmax = top + 1 > kmax ? top + 1 : kmax;
    }
    kmax = max(kmax, pair_swap(kmax));
    f.exp2 = emin;
    f.shifted = kmax > kWideCoordBits;
    const i128 x0O[3] = {pair_swap(X[0]), pair_swap(X[1]), pair_swap(X[2])};
#pragma unroll
    for (int k = 0; k < 3; ++k) f.x0[k] = f.shifted ? (sideA ? X[k] : x0O[k]) : (i128)0;
    int k1 = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const i128 d = X[i] - f.x0[i % 3];
        const int bt = bits128(d);
        k1 = bt > k1 ? bt : k1;
        T[i / 3].v[i % 3] = d;
    }
    k1 = max(k1, pair_swap(k1));
    ok = ok & pair_swap(ok);
    return ok && k1 <= kWideCoordBits;
}

__global__ __launch_bounds__(256) void k_exact_wide_fast(MeshDev A, MeshDev B, const uint2* __restrict__ list,
                                                         uint32_t cap, const uint32_t* __restrict__ grid_flags,
                                                         int bitsB, uint64_t* __restrict__ rec_key,
                                                         uint2* __restrict__ rec_seg, uint32_t rec_cap,
                                                         uint64_t* __restrict__ seg_pool, int32_t* __restrict__ seg_exp,
                                                         uint32_t seg_cap, uint2* __restrict__ slow, uint32_t slow_cap,
                                                         uint2* __restrict__ resid, uint32_t resid_cap,
                                                         uint32_t* __restrict__ counters) {
    if (grid_flags[0] && grid_flags[1]) return;  // grid meshes: k_exact_grid_fast
    uint32_t n = counters[CNT_EXACT_LIST];
    if (n > cap) n = cap;
    ExactTally tally;
    // two lanes per pair: the even lane holds triangle A and writes the results
    const bool sideA = (threadIdx.x & 1u) == 0u;
    for (uint32_t base = blockIdx.x * 128; base < n; base += gridDim.x * 128) {
        const uint32_t i = base + (threadIdx.x >> 1);
        bool residual = false, to_slow = false, hit = false;
        uint2 en = make_uint2(0, 0);
        WideExact::HP lo, hi;
        WideFrame fr;
        fr.exp2 = 0;
        fr.shifted = 0;
        if (i < n) {
            en = list[i];
            const uint32_t* tr = sideA ? A.tri + 3 * (size_t)en.x : B.tri + 3 * (size_t)(en.y & kIndexMask);
            const double4* v64 = sideA ? A.v64 : B.v64;
            const double* src = sideA ? A.src : B.src;
            double t[9];
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                const uint32_t j = tr[v];
                if (v64) {
                    const double4 p = v64[j];
                    t[3 * v] = p.x; t[3 * v + 1] = p.y; t[3 * v + 2] = p.z;
                } else {
                    const double* p = src + 3 * (size_t)j;
                    t[3 * v] = p[0]; t[3 * v + 1] = p[1]; t[3 * v + 2] = p[2];
                }
            }
            WideExact::V3 T[3];
            residual = !wide_frame_pair(t, sideA, T, fr);
            if (!residual) {
                uint32_t np = 0;
                const int r = transversal_segment_pair<WideExact, WidePolicy>(T, lo, hi, np);
                to_slow = r < 0;
                hit = r > 0;
                if (hit && sideA) residual = !wide_absolute(fr, lo) || !wide_absolute(fr, hi);
                if (residual) {
                    hit = false;
                } else if (sideA) {
                    tally.pred += np;
                    if (!to_slow) tally.add((int)(en.y >> kStatusShift), hit, 0, 0);
                }
            }
            residual = residual && sideA;
            to_slow = to_slow && sideA;
            hit = hit && sideA;
        }
        const uint32_t rslot = block_reserve(residual ? 1u : 0u, counters, CNT_RESIDUAL);
        if (residual && rslot < resid_cap) resid[rslot] = en;
        const uint32_t want[3] = {to_slow ? 1u : 0u, hit ? 1u : 0u, hit ? 1u : 0u};
        uint32_t got[3];
        block_reserve3(want, counters, CNT_SLOW, CNT_RESULTS, CNT_SEGMENTS, got);
        if (to_slow && got[0] < slow_cap) slow[got[0]] = en;
        if (hit && got[1] < rec_cap) {
            rec_key[got[1]] = ((uint64_t)en.x << bitsB) | (uint64_t)(en.y & kIndexMask);
            rec_seg[got[1]] = make_uint2(got[2], 1u);
            if (got[2] < seg_cap) {
                uint64_t* o = seg_pool + (size_t)8 * kSegLimbsWide * got[2];
                store_w256(o + 0, lo.h[0]); store_w256(o + 4, lo.h[1]); store_w256(o + 8, lo.h[2]);
                store_w256(o + 12, lo.w);
                store_w256(o + 16, hi.h[0]); store_w256(o + 20, hi.h[1]); store_w256(o + 24, hi.h[2]);
                store_w256(o + 28, hi.w);
                seg_exp[got[2]] = fr.exp2;
            }
        }
    }
    tally.flush(counters);
}

// The general doubles' pairs k_exact_wide_fast left in `slow` (CNT_SLOW):
// the general one-lane pass, then its deferred coplanar pairs.
void launch_exact_wide_slow(MeshDev A, MeshDev B, const uint32_t* grid_flags, int bitsB, uint64_t* rec_key,
                            uint2* rec_seg, uint32_t rec_cap, uint64_t* seg_pool, int32_t* seg_exp, uint32_t seg_cap,
                            uint2* defer, uint32_t defer_cap, uint2* slow, uint32_t slow_cap, uint2* resid,
                            uint32_t resid_cap, uint32_t* counters, int blocks, hipStream_t s) {
    if (A.nt == 0 || B.nt == 0) return;
    k_exact_wide<true><<<blocks, 256, 0, s>>>(A, B, slow, slow_cap, grid_flags, bitsB, rec_key, rec_seg, rec_cap,
                                              seg_pool, seg_exp, seg_cap, defer, defer_cap, resid, resid_cap, counters,
                                              (int)CNT_SLOW);
    k_exact_wide<false><<<blocks, 256, 0, s>>>(A, B, defer, defer_cap, grid_flags, bitsB, rec_key, rec_seg, rec_cap,
                                               seg_pool, seg_exp, seg_cap, nullptr, 0, resid, resid_cap, counters,
                                               (int)CNT_DEFERRED);
    GI_LAUNCH_CHECK();
}

// The grid pairs the one-lane pass left in `slow` (counted in CNT_SLOW): the
// 8-lane pass, then the deferred (coplanar, degenerate) pairs it hands on,
// one lane each with the full construction (their count is read on the
// device; usually few: a quarter of the blocks).
void launch_exact_grid_slow(MeshDev A, MeshDev B, const uint32_t* grid_flags, int bitsB, uint64_t* rec_key,
                            uint2* rec_seg, uint32_t rec_cap, uint64_t* seg_pool, uint32_t seg_cap, uint2* defer,
                            uint32_t defer_cap, uint2* slow, uint32_t slow_cap, uint32_t* counters, int blocks,
                            hipStream_t s) {
    if (A.nt == 0 || B.nt == 0) return;
    k_exact_grid_lanes<<<blocks, 256, 0, s>>>(A, B, slow, slow_cap, grid_flags, bitsB, rec_key, rec_seg, rec_cap,
                                              seg_pool, seg_cap, defer, defer_cap, counters, (int)CNT_SLOW);
    k_exact<false><<<std::max(blocks / 4, 64), 256, 0, s>>>(A, B, defer, defer_cap, grid_flags, bitsB, rec_key, rec_seg,
                                                            rec_cap, seg_pool, seg_cap, nullptr, 0, counters);
    GI_LAUNCH_CHECK();
}

void launch_exact(MeshDev A, MeshDev B, const uint2* exact_list, uint32_t exact_cap, const uint32_t* grid_flags,
                  int bitsB, uint64_t* rec_key, uint2* rec_seg, uint32_t rec_cap, uint64_t* seg_pool,
                  int32_t* seg_exp, uint32_t seg_cap, uint2* defer, uint32_t defer_cap, uint2* slow,
                  uint32_t slow_cap, uint2* resid, uint32_t resid_cap, uint32_t* counters, int blocks, int kinds,
                  hipStream_t s) {
    if (A.nt == 0 || B.nt == 0) return;
    // kinds & 1: the grid stage, & 2: the wide stage; with both, the one the
    // grid flags (decided on the device by the vertex stage) do not select
    // returns at once
    // pairs one lane each; the coplanar and degenerate ones to the slow
    // list, whose passes run here or, with
    // kExactGridSlowLater, when the caller has read CNT_SLOW
    const bool later = (kinds & kExactGridSlowLater) != 0;
    if (kinds & kExactGrid) {
        k_exact_grid_fast<<<blocks, 256, 0, s>>>(A, B, exact_list, exact_cap, grid_flags, bitsB, rec_key, rec_seg,
                                                 rec_cap, seg_pool, seg_cap, slow, slow_cap, counters);
        if (!later)
            launch_exact_grid_slow(A, B, grid_flags, bitsB, rec_key, rec_seg, rec_cap, seg_pool, seg_cap, defer,
                                   defer_cap, slow, slow_cap, counters, blocks, s);
    }
    if (kinds & kExactWide) {
        k_exact_wide_fast<<<blocks, 256, 0, s>>>(A, B, exact_list, exact_cap, grid_flags, bitsB, rec_key, rec_seg,
                                                 rec_cap, seg_pool, seg_exp, seg_cap, slow, slow_cap, resid, resid_cap,
                                                 counters);
        if (!later)
            launch_exact_wide_slow(A, B, grid_flags, bitsB, rec_key, rec_seg, rec_cap, seg_pool, seg_exp, seg_cap,
                                   defer, defer_cap, slow, slow_cap, resid, resid_cap, counters, blocks, s);
    }
    GI_LAUNCH_CHECK();
}

// ----------------------------------------------------------------- gathers

__global__ void k_iota(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

void launch_iota(uint32_t* p, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    k_iota<<<blocks_for(n, 256), 256, 0, s>>>(p, n);
    GI_LAUNCH_CHECK();
}

__global__ void k_gather_pairs(const uint64_t* __restrict__ key, const uint32_t* __restrict__ order,
                               const uint2* __restrict__ rec_seg, uint32_t n, int bitsB, uint2* __restrict__ pairs,
                               uint32_t* __restrict__ nseg) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = key[i];
    pairs[i] = make_uint2((uint32_t)(k >> bitsB), (uint32_t)(k & ((1ull << bitsB) - 1ull)));
    nseg[i] = rec_seg[order[i]].y;
}

void launch_gather_pairs(const uint64_t* sorted_key, const uint32_t* order, const uint2* rec_seg, uint32_t n, int bitsB,
                         uint2* pairs, uint32_t* nseg, hipStream_t s) {
    if (n == 0) return;
    k_gather_pairs<<<blocks_for(n, 256), 256, 0, s>>>(sorted_key, order, rec_seg, n, bitsB, pairs, nseg);
    GI_LAUNCH_CHECK();
}

// an L-word (2 or 4) two's complement integer as a double: its magnitude's
// words summed from the top, the sign applied last
template <int L>
__device__ __forceinline__ double limbs_to_double(const uint64_t* __restrict__ x) {
    const bool neg = (x[L - 1] >> 63) != 0;
    uint64_t m[L];
    uint64_t c = 1;
#pragma unroll
    for (int l = 0; l < L; ++l) {
        const uint64_t v = neg ? ~x[l] : x[l];
        m[l] = neg ? v + c : v;
        c = neg ? (c && m[l] == 0) : 0;
    }
    double d = 0.0;
#pragma unroll
    for (int l = L - 1; l >= 0; --l) d = d * 18446744073709551616.0 + (double)m[l];
    return neg ? -d : d;
}

// the rs.y segments of one record (at rs.x in the pool) to dst onward:
// homogeneous endpoints of L words, their power-of-two scales (wide form,
// pool_exp non-null), and the doubles they divide to
template <int L>
__device__ __forceinline__ void copy_record_segments(uint2 rs, uint32_t dst, const uint64_t* __restrict__ pool,
                                                     const int32_t* __restrict__ pool_exp,
                                                     uint64_t* __restrict__ out, int32_t* __restrict__ out_exp,
                                                     double* __restrict__ f64) {
    constexpr int kWords = 8 * L;
    for (uint32_t s = 0; s < rs.y; ++s) {
        const uint64_t* src = pool + (size_t)kWords * (rs.x + s);
        uint64_t* o = out + (size_t)kWords * (dst + s);
        double* f = f64 + 6 * (size_t)(dst + s);
        uint64_t w[kWords];
#pragma unroll
        for (int i = 0; i < kWords; i += 2) {
            const ulonglong2 v = *(const ulonglong2*)(src + i);
            w[i] = v.x;
            w[i + 1] = v.y;
        }
#pragma unroll
        for (int i = 0; i < kWords; i += 2) *(ulonglong2*)(o + i) = make_ulonglong2(w[i], w[i + 1]);
        const int e2 = pool_exp ? pool_exp[rs.x + s] : 0;
        if (out_exp) out_exp[dst + s] = e2;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const double wd = limbs_to_double<L>(w + (4 * e + 3) * L);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double v = limbs_to_double<L>(w + (4 * e + k) * L) / wd;
                f[3 * e + k] = pool_exp ? ldexp(v, e2) : v;
            }
        }
    }
}

__global__ void k_gather_segments(const uint32_t* __restrict__ order, const uint2* __restrict__ rec_seg,
                                  const uint32_t* __restrict__ seg_offsets, uint32_t n,
                                  const uint64_t* __restrict__ pool, const int32_t* __restrict__ pool_exp, int limbs,
                                  uint64_t* __restrict__ out, int32_t* __restrict__ out_exp,
                                  double* __restrict__ f64) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (limbs == kSegLimbsGrid) copy_record_segments<kSegLimbsGrid>(rec_seg[order[i]], seg_offsets[i], pool, nullptr,
                                                                     out, nullptr, f64);
    else copy_record_segments<kSegLimbsWide>(rec_seg[order[i]], seg_offsets[i], pool, pool_exp, out, out_exp, f64);
}

void launch_gather_segments(const uint32_t* order, const uint2* rec_seg, const uint32_t* seg_offsets, uint32_t n,
                            const uint64_t* seg_pool, const int32_t* seg_exp, int limbs, uint64_t* seg_out,
                            int32_t* seg_exp_out, double* seg_f64, hipStream_t s) {
    if (n == 0) return;
    k_gather_segments<<<blocks_for(n, 256), 256, 0, s>>>(order, rec_seg, seg_offsets, n, seg_pool,
                                                         limbs == kSegLimbsWide ? seg_exp : nullptr, limbs, seg_out,
                                                         limbs == kSegLimbsWide ? seg_exp_out : nullptr, seg_f64);
    GI_LAUNCH_CHECK();
}

// Sorted output of up to kOutSmallMax result records in one launch, queued
// before the host reads the counters: the record count R comes from device
// memory, and nothing runs unless R fits (the host then takes the multi-tile
// path after the read-back, 20-odd launches). Each key (faceA << bitsB |
// faceB) carries its record index in the low idx_bits, so the sort moves no
// separate values (keys are distinct).
constexpr uint32_t kOutSmallMax = 16384;

__host__ __device__ __forceinline__ int bits_of_count(uint32_t n) { return n <= 1 ? 1 : 32 - __builtin_clz(n - 1); }

// the records and segments were all stored (no overflow) and fit
__device__ __forceinline__ bool output_small_ok(const uint32_t* counters, uint32_t cap_rec, uint32_t cap_seg,
                                                int key_bits, uint32_t& R) {
    R = counters[CNT_RESULTS];
    return R <= cap_rec && counters[CNT_SEGMENTS] <= cap_seg && R <= kOutSmallMax &&
           key_bits + bits_of_count(R) <= 63;
}

// 64-bit value of lane (lane ^ M) of the wave: DPP quad permutes for
// M = 1, 2, LDS-free swizzles within 32 lanes for M = 4 .. 16, a permute
// through the LDS crossbar for 32.
template <int M>
__device__ __forceinline__ uint64_t lane_xor_u64(uint64_t x) {
    int lo = (int)(uint32_t)x, hi = (int)(uint32_t)(x >> 32);
    if constexpr (M == 1) {
        lo = __builtin_amdgcn_update_dpp(0, lo, 0xB1, 0xF, 0xF, false);  // quad_perm [1, 0, 3, 2]
        hi = __builtin_amdgcn_update_dpp(0, hi, 0xB1, 0xF, 0xF, false);
    } else if constexpr (M == 2) {
        lo = __builtin_amdgcn_update_dpp(0, lo, 0x4E, 0xF, 0xF, false);  // quad_perm [2, 3, 0, 1]
        hi = __builtin_amdgcn_update_dpp(0, hi, 0x4E, 0xF, 0xF, false);
    } else if constexpr (M < 32) {
        lo = __builtin_amdgcn_ds_swizzle(lo, (M << 10) | 0x1F);  // bit mode: lane ^ M within 32
        hi = __builtin_amdgcn_ds_swizzle(hi, (M << 10) | 0x1F);
    } else {
        lo = __shfl_xor(lo, M);
        hi = __shfl_xor(hi, M);
    }
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Bitonic sort of a block's keys, thread t holding positions 2t and 2t + 1
// (k0, k1). Merge K, step J compares positions i and i ^ J, keeping
// the smaller one at the lower position where i & K is zero (ascending
// blocks) and the larger one elsewhere. J = 1 is within a thread, J = 2 ..
// 64 between lanes of a wave (lane distance J / 2), J >= 128 between waves
// through LDS (s), one pair per thread.
template <int K, int J>
__device__ __forceinline__ void run_bitonic_step(uint64_t& k0, uint64_t& k1, uint64_t* s, uint32_t t, bool active) {
    if constexpr (J >= 128) {
        const uint32_t i = ((t & ~(uint32_t)(J - 1)) << 1) | (t & (uint32_t)(J - 1));
        const bool asc = (i & (uint32_t)K) == 0;
        const uint64_t a = s[i], b = s[i + J];
        const bool swap = asc ? a > b : a < b;
        s[i] = swap ? b : a;
        s[i + J] = swap ? a : b;
        __syncthreads();
    } else if constexpr (J >= 2) {
        if (!active) return;  // the whole wave (active is the same on its lanes)
        const uint64_t p0 = lane_xor_u64<J / 2>(k0), p1 = lane_xor_u64<J / 2>(k1);
        const bool keep_min = ((t & (uint32_t)(J / 2)) == 0) == ((t & (uint32_t)(K / 2)) == 0);
        k0 = keep_min ? (k0 < p0 ? k0 : p0) : (k0 < p0 ? p0 : k0);
        k1 = keep_min ? (k1 < p1 ? k1 : p1) : (k1 < p1 ? p1 : k1);
    } else {
        if (!active) return;
        const bool asc = (t & (uint32_t)(K / 2)) == 0;
        const uint64_t lo = k0 < k1 ? k0 : k1, hi = k0 < k1 ? k1 : k0;
        k0 = asc ? lo : hi;
        k1 = asc ? hi : lo;
    }
}
template <int K, int J>
struct RunBitonicMerge {
    __device__ static __forceinline__ void run(uint64_t& k0, uint64_t& k1, uint64_t* s, uint32_t t, bool active) {
        if constexpr (J == K / 2 && J >= 128) {  // into LDS for the cross-wave steps
            s[2 * t] = k0;
            s[2 * t + 1] = k1;
            __syncthreads();
        }
        run_bitonic_step<K, J>(k0, k1, s, t, active);
        if constexpr (J == 128) {  // back to registers
            k0 = s[2 * t];
            k1 = s[2 * t + 1];
            __syncthreads();
        }
        RunBitonicMerge<K, J / 2>::run(k0, k1, s, t, active);
    }
};
template <int K>
struct RunBitonicMerge<K, 0> {
    __device__ static __forceinline__ void run(uint64_t&, uint64_t&, uint64_t*, uint32_t, bool) {}
};
// Every aligned block of K positions of the block's 2 x blockDim keys ends
// sorted (the first one ascending). Waves with active false skip the
// steps within registers; all threads take part in the LDS steps (K >= 256).
template <int K>
struct RunBitonicSort {
    __device__ static __forceinline__ void run(uint64_t& k0, uint64_t& k1, uint64_t* s, uint32_t t, bool active) {
        RunBitonicSort<K / 2>::run(k0, k1, s, t, active);
        RunBitonicMerge<K, K / 2>::run(k0, k1, s, t, active);
    }
};
template <>
struct RunBitonicSort<1> {
    __device__ static __forceinline__ void run(uint64_t&, uint64_t&, uint64_t*, uint32_t, bool) {}
};

// segments of the sorted records, for the records k_output_buckets ordered
__global__ __launch_bounds__(256) void k_gather_segments_small(const uint32_t* __restrict__ order,
                                                               const uint2* __restrict__ rec_seg,
                                                               const uint32_t* __restrict__ seg_offsets,
                                                               const uint32_t* __restrict__ counters, uint32_t cap_rec,
                                                               uint32_t cap_seg, int key_bits, uint32_t out_slots,
                                                               const uint64_t* __restrict__ pool,
                                                               const int32_t* __restrict__ pool_exp,
                                                               const uint32_t* __restrict__ grid_flags,
                                                               uint64_t* __restrict__ out,
                                                               int32_t* __restrict__ out_exp,
                                                               double* __restrict__ f64) {
    uint32_t R;
    if (!output_small_ok(counters, cap_rec, cap_seg, key_bits, R) || seg_offsets[R] > out_slots) return;
    const bool grid = grid_flags[0] && grid_flags[1];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < R; i += gridDim.x * blockDim.x)
        if (grid) copy_record_segments<kSegLimbsGrid>(rec_seg[order[i]], seg_offsets[i], pool, nullptr, out, nullptr, f64);
        else copy_record_segments<kSegLimbsWide>(rec_seg[order[i]], seg_offsets[i], pool, pool_exp, out, out_exp, f64);
}

// kOutBuckets workgroups, each owning the records whose key has its index
// in the key's top eight bits (a range of A's faces). Every workgroup reads
// all R records: it counts those of lower ranges and their segments (its
// output starts there) and packs its own into LDS in record order; a
// bitonic sort in registers orders them (up to 2 x OB_THREADS keys; the
// waves past the keys skip the in-register steps, and up to 128 keys take
// no barrier), or in LDS for a larger range (any size up to R); then each
// goes straight to its place with its pair, record index and segment offset
// (a block scan of the segment counts in sorted order). Measured against
// 64 ranges sorted in LDS with a block barrier per step (19.7 us for 10^4
// records) and two launches (2,048-record runs, then every record's rank in
// the other runs by binary searches: 24 us).
constexpr uint32_t kOutBuckets = 256;
constexpr int OB_THREADS = 1024, OB_ITEMS = (int)(kOutSmallMax / OB_THREADS);
static_assert(kOutSmallMax % OB_THREADS == 0, "whole items per thread");

// Block-wide sums over the 1024-thread block of four counts at once: the
// exclusive prefix of a and the totals of all four.
__device__ __forceinline__ uint32_t ob_scan4(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t (&tot)[4],
                                            uint4* ws) {
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    uint32_t x = a;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= (uint32_t)o) x += y;
    }
    for (int o = 32; o > 0; o >>= 1) {
        b += __shfl_xor(b, o);
        c += __shfl_xor(c, o);
        d += __shfl_xor(d, o);
    }
    if (lane == 63) ws[wave] = make_uint4(x, b, c, d);
    __syncthreads();
    uint32_t off = 0;
    tot[0] = tot[1] = tot[2] = tot[3] = 0;
    for (uint32_t w = 0; w < OB_THREADS / 64; ++w) {
        const uint4 q = ws[w];
        off += w < wave ? q.x : 0u;
        tot[0] += q.x;
        tot[1] += q.y;
        tot[2] += q.z;
        tot[3] += q.w;
    }
    __syncthreads();
    return off + x - a;
}

// exclusive sum over the 1024-thread block; total to tot
__device__ __forceinline__ uint32_t ob_exclusive_sum(uint32_t v, uint32_t& tot, uint4* ws) {
    uint32_t t4[4];
    const uint32_t e = ob_scan4(v, 0u, 0u, 0u, t4, ws);
    tot = t4[0];
    return e;
}

template <int K>
__device__ __forceinline__ void ob_sort_regs(uint64_t* sk, uint32_t t) {
    uint64_t k0 = sk[2 * t], k1 = sk[2 * t + 1];
    if constexpr (K >= 256) __syncthreads();  // every key read before the LDS steps write sk
    RunBitonicSort<K>::run(k0, k1, sk, t, (t & ~63u) * 2 < (uint32_t)K);  // the wave holds keys of the block
    if (2 * t < (uint32_t)K) {
        sk[2 * t] = k0;
        sk[2 * t + 1] = k1;
    }
}

__global__ __launch_bounds__(OB_THREADS) void k_output_buckets(const uint64_t* __restrict__ rec_key,
                                                               const uint2* __restrict__ rec_seg,
                                                               const uint32_t* __restrict__ counters, uint32_t cap_rec,
                                                               uint32_t cap_seg, int key_bits, int bitsB,
                                                               int want_segments, uint2* __restrict__ pairs,
                                                               uint32_t* __restrict__ order,
                                                               uint32_t* __restrict__ offs,
                                                               uint32_t* __restrict__ counters_out) {
    // the counters are final here (every stage that counts has run): the
    // first workgroup copies the counter words (one per kCntStride) and the
    // two grid flags to the host's pinned read-back buffer
    if (blockIdx.x == 0 && counters_out && threadIdx.x < CNT_COUNT / kCntStride + 2) {
        const uint32_t i = threadIdx.x < CNT_COUNT / kCntStride ? threadIdx.x * kCntStride
                                                                 : CNT_COUNT + threadIdx.x - CNT_COUNT / kCntStride;
        counters_out[i] = counters[i];
    }
    uint32_t R;
    if (!output_small_ok(counters, cap_rec, cap_seg, key_bits, R)) return;
    __shared__ uint64_t sk[kOutSmallMax];
    __shared__ uint4 ws[OB_THREADS / 64];
    const uint32_t t = threadIdx.x, mine = blockIdx.x;
    const int idx_bits = bits_of_count(R);
    const int sh = key_bits > 8 ? key_bits - 8 : 0;  // the record keys are below 2^key_bits
    // this thread's records: t, t + OB_THREADS, ... (all loaded before any
    // is looked at)
    uint64_t rk[OB_ITEMS];
    uint32_t ns[OB_ITEMS];
#pragma unroll
    for (int it = 0; it < OB_ITEMS; ++it) {
        const uint32_t i = it * OB_THREADS + t;
        rk[it] = i < R ? rec_key[i] : 0ull;
        ns[it] = (i < R && want_segments) ? rec_seg[i].y : 0u;
    }
    uint32_t own = 0, below = 0, seg_below = 0, seg_all = 0, own_mask = 0;
#pragma unroll
    for (int it = 0; it < OB_ITEMS; ++it) {
        if (it * OB_THREADS + t >= R) continue;
        const uint32_t b = (uint32_t)(rk[it] >> sh);
        seg_all += ns[it];
        below += b < mine ? 1u : 0u;
        seg_below += b < mine ? ns[it] : 0u;
        if (b == mine) {
            own_mask |= 1u << it;
            ++own;
        }
    }
    uint32_t tot[4];
    uint32_t pos = ob_scan4(own, below, seg_below, seg_all, tot, ws);
    const uint32_t n = tot[0];
    below = tot[1];
    seg_below = tot[2];
    if (mine == 0 && t == 0) offs[R] = tot[3];
    if (n == 0) return;
#pragma unroll
    for (int it = 0; it < OB_ITEMS; ++it)
        if ((own_mask >> it) & 1u) sk[pos++] = (rk[it] << idx_bits) | (it * OB_THREADS + t);
    uint32_t m = 2;
    while (m < n) m <<= 1;
    for (uint32_t i = n + t; i < m; i += OB_THREADS) sk[i] = ~0ull;  // real keys stay below 2^63
    __syncthreads();
    if (m <= 2u * OB_THREADS) {
        switch (m) {  // the same on every thread
            case 2: if (t < 64) ob_sort_regs<2>(sk, t); break;
            case 4: if (t < 64) ob_sort_regs<4>(sk, t); break;
            case 8: if (t < 64) ob_sort_regs<8>(sk, t); break;
            case 16: if (t < 64) ob_sort_regs<16>(sk, t); break;
            case 32: if (t < 64) ob_sort_regs<32>(sk, t); break;
            case 64: if (t < 64) ob_sort_regs<64>(sk, t); break;
            case 128: if (t < 64) ob_sort_regs<128>(sk, t); break;
            case 256: ob_sort_regs<256>(sk, t); break;
            case 512: ob_sort_regs<512>(sk, t); break;
            case 1024: ob_sort_regs<1024>(sk, t); break;
            default: ob_sort_regs<2048>(sk, t); break;
        }
    } else {  // in LDS, a block barrier per step
        for (uint32_t k = 2; k <= m; k <<= 1)
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t i = t; i < m / 2; i += OB_THREADS) {
                    const uint32_t lo = ((i & ~(j - 1u)) << 1) | (i & (j - 1u)), hi = lo + j;
                    const uint64_t a = sk[lo], b = sk[hi];
                    if ((a > b) == ((lo & k) == 0u)) {
                        sk[lo] = b;
                        sk[hi] = a;
                    }
                }
                __syncthreads();
            }
    }
    __syncthreads();
    const uint64_t idx_mask = (1ull << idx_bits) - 1ull, maskB = (1ull << bitsB) - 1ull;
    uint32_t carry = seg_below;
    for (uint32_t p0 = 0; p0 < n; p0 += OB_THREADS) {  // the same rounds on every thread
        const uint32_t p = p0 + t;
        const uint64_t x = p < n ? sk[p] : 0ull;
        const uint32_t rec = (uint32_t)(x & idx_mask);
        const uint32_t nsp = (p < n && want_segments) ? rec_seg[rec].y : 0u;
        uint32_t all;
        const uint32_t e = ob_exclusive_sum(nsp, all, ws);
        if (p < n) {
            const uint64_t pk = x >> idx_bits;
            pairs[below + p] = make_uint2((uint32_t)(pk >> bitsB), (uint32_t)(pk & maskB));
            order[below + p] = rec;
            offs[below + p] = carry + e;
        }
        carry += all;
    }
}

uint32_t output_small_max() { return kOutSmallMax; }

bool output_small_fits(uint32_t R, int key_bits) { return R <= kOutSmallMax && key_bits + bits_of_count(R) <= 63; }

void launch_output_small(const uint64_t* rec_key, const uint2* rec_seg, const uint32_t* counters, uint32_t cap_rec,
                         uint32_t cap_seg, int key_bits, int bitsB, bool want_segments, uint2* pairs,
                         uint32_t* order, uint32_t* offs, uint64_t* seg_out, int32_t* seg_exp_out, double* seg_f64,
                         uint32_t out_slots, const uint64_t* seg_pool, const int32_t* seg_exp,
                         const uint32_t* grid_flags, uint32_t* counters_out, hipStream_t s) {
    k_output_buckets<<<kOutBuckets, OB_THREADS, 0, s>>>(rec_key, rec_seg, counters, cap_rec, cap_seg, key_bits, bitsB,
                                                        want_segments ? 1 : 0, pairs, order, offs, counters_out);
    if (want_segments)
        k_gather_segments_small<<<kOutSmallMax / 256, 256, 0, s>>>(order, rec_seg, offs, counters, cap_rec, cap_seg,
                                                                   key_bits, out_slots, seg_pool, seg_exp, grid_flags,
                                                                   seg_out, seg_exp_out, seg_f64);
    GI_LAUNCH_CHECK();
}

// ------------------------------------------------------ forest results
//
// The sorted pairs of a forest are ordered by global A index, so mesh m's
// pairs start at the first pair whose A triangle is at or after toffA[m].

__global__ void k_pair_offsets(const uint2* __restrict__ pairs, uint32_t R, const uint32_t* __restrict__ toffA,
                               uint32_t M, uint32_t* __restrict__ pair_off) {
    uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m > M) return;
    uint32_t key = toffA[m], lo = 0, hi = R;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (pairs[mid].x < key) lo = mid + 1;
        else hi = mid;
    }
    pair_off[m] = m == M ? R : lo;
}

// global -> per-mesh triangle indices: pair i belongs to the last mesh m
// with pair_off[m] <= i
__global__ void k_localize(uint2* __restrict__ pairs, uint32_t R, const uint32_t* __restrict__ pair_off, uint32_t M,
                           const uint32_t* __restrict__ toffA, const uint32_t* __restrict__ toffB) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    uint32_t lo = 0, hi = M;  // invariant: pair_off[lo] <= i < pair_off[hi]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (pair_off[mid] <= i) lo = mid;
        else hi = mid;
    }
    uint2 p = pairs[i];
    pairs[i] = make_uint2(p.x - toffA[lo], p.y - toffB[lo]);
}

void launch_forest_localize(uint2* pairs, uint32_t R, const uint32_t* toffA, const uint32_t* toffB, uint32_t M,
                            uint32_t* pair_off, hipStream_t s) {
    k_pair_offsets<<<blocks_for((size_t)M + 1, 256), 256, 0, s>>>(pairs, R, toffA, M, pair_off);
    if (R) k_localize<<<blocks_for(R, 256), 256, 0, s>>>(pairs, R, pair_off, M, toffA, toffB);
    GI_LAUNCH_CHECK();
}

}  // namespace gi
