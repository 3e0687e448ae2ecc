/*
 * gi_intersector.h -- C ABI of the MI355X mesh/mesh intersection library
 * (libgi.so). Plain C types only; every function returns a gi_status and
 * never exits the process.
 *
 * The entry points replace the reference's GPUIntersector surface
 * (paths relative to the reference root):
 *
 *   gi_create / gi_destroy      KernelBVHController::KernelBVHController,
 *                               ~KernelBVHController, cleanup
 *                               (src/GPUIntersector/KernelBVHController.h:38-39, 66)
 *   gi_set_mesh                 the per-mesh half of KernelBVHController::construct
 *                               (KernelBVHController.h:46-57; .cu:444-597)
 *   gi_set_transform            KernelBVHController::setTransformBoth /
 *                               setTranslation (KernelBVHController.h:72-80;
 *                               .cu:289-412), one mesh per call
 *   gi_build                    the BVH build inside construct (.cu:546-580)
 *   gi_intersect                KernelBVHController::runIntersectionPipeline
 *                               (KernelBVHController.h:60-63; .cu:603-699) fused
 *                               with computeExactIntersectionCurves
 *                               (src/CPU/MeshIntersectionCurves.h:24-29)
 *   gi_intersect_oneshot        extern "C" runGridForestIntersectionPipeline
 *                               (src/GPUIntersector/StandaloneBVHPipeline.h:27-42):
 *                               everything built per call from host arrays
 *   gi_clear_device /           KernelBVHController::clearGPU / reconstructGPU
 *   gi_reconstruct_device       (KernelBVHController.h:67-69; .cu:731-948)
 *   gi_result_free              the caller's std::free of the pair buffer
 *                               (apps/mainApplication.cpp:793-796)
 *   gi_exact_pair_host          CGAL::do_intersect + CGAL::intersection on one
 *                               triangle pair (src/CPU/YellowFilter.h:122,
 *                               src/CPU/MeshIntersectionCurves.cpp:63), as run by
 *                               the host residual stage
 *
 * Semantics shared by all entry points:
 *   - triangles are closed point sets; a zero-area triangle is the segment or
 *     point it spans;
 *   - a pair (a, b) is reported iff triangle a of mesh A and triangle b of
 *     mesh B have a common point, decided exactly on the double coordinates
 *     after the rigid transform;
 *   - pairs come back sorted by (a, b) and unique;
 *   - segments are exact: non-coplanar pairs give one segment (a point is a
 *     segment with equal endpoints), coplanar pairs give the boundary edges of
 *     their overlap polygon.
 */
#ifndef GI_INTERSECTOR_H
#define GI_INTERSECTOR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    GI_OK = 0,
    GI_ERR_INVALID_ARGUMENT = 1,
    GI_ERR_HIP = 2,
    GI_ERR_OUT_OF_MEMORY = 3,
    GI_ERR_NO_DEVICE = 4,
    GI_ERR_NOT_READY = 5,   /* a mesh slot has not been set */
    GI_ERR_INTERNAL = 6,    /* an internal consistency check failed */
    GI_ERR_TOO_LARGE = 7,   /* more than 2^30 triangles in one mesh */
    GI_ERR_IO = 8           /* file cannot be read, parsed or written (gi_meshio.h) */
} gi_status;

typedef struct gi_context gi_context;

typedef struct {
    int32_t device;             /* HIP device ordinal */
    uint32_t candidate_reserve; /* initial candidate-queue capacity, 0 = auto */
    int32_t host_threads;       /* threads of the host residual stage, 0 = auto */
    int32_t reserved;
} gi_config;

/* slot 0 is mesh A, slot 1 is mesh B */
enum { GI_MESH_A = 0, GI_MESH_B = 1 };

typedef struct {
    int32_t output_on_device;   /* 1: result arrays stay in device memory, owned by
                                   the context until its next gi_intersect */
    int32_t want_segments;      /* 0: pairs only */
    int32_t rebuild;            /* 1: rebuild both BVHs from scratch before the query */
    int32_t use_graph;          /* 1: replay the query's launches up to the counter read-back as a
                                   HIP graph while sizes and buffers stay the same (captured on
                                   the second such query; its timing events are graph nodes) */
    int32_t stage_timing;       /* HIP events that time the stages (each costs the queue a few
                                   microseconds): 0 every stage (all ms_* statistics); otherwise
                                   only the launches in the mask GI_TIME_STAGING | GI_TIME_BVH_CHUNKS
                                   | GI_TIME_TRAVERSE (GI_TIME_NONE: no event at all) */
    int32_t reserved;
} gi_options;

enum {
    GI_TIME_STAGING = 1,    /* the staging launch of both meshes (ms_stage_meshes) */
    GI_TIME_BVH_CHUNKS = 2, /* the k_build_chunks launch (ms_bvh_chunks) */
    GI_TIME_TRAVERSE = 4,   /* the traversal (ms_traverse) */
    GI_TIME_NONE = 8
};

typedef struct {
    /* candidate pipeline: traversal (FP32 boxes) -> exact boxes -> filter -> exact */
    uint64_t traversal_hits;    /* FP32 box overlaps found by the BVH traversal */
    uint64_t candidates;        /* exact closed double-box overlaps */
    uint64_t filter_no, filter_green, filter_yellow; /* FP64 filter outcomes */
    uint64_t exact_pairs;       /* pairs run through the exact integer stage (GPU) */
    uint64_t exact_yellow_yes, exact_yellow_no;      /* YELLOW pairs resolved exactly */
    uint64_t host_pairs;        /* residual pairs run through the host stage (general doubles whose
                                   integer frame exceeds the GPU stage's width) */
    uint64_t host_yes;
    uint64_t pairs, segments, coplanar, degenerate;
    /* predicates evaluated per cascade stage (orient3d signs in the FP64
       filter; orient3d and orient2d signs in the exact stages) */
    uint64_t filter_predicates, exact_predicates, host_predicates;
    int32_t grid;               /* 1: both meshes on the integer grid (int64/int128 exact stage);
                                   0: general doubles (per-pair integer frames, 256-bit) */
    int32_t retries;            /* capacity regrowths during this call */
    /* milliseconds (GPU events; host stage by wall clock) */
    float ms_total, ms_build, ms_traverse, ms_filter, ms_exact, ms_output, ms_host;
    float ms_bvh_chunks;        /* the k_build_chunks launch of this call (A's tree; a B shared into an A slot builds in the same launch) */
    float ms_stage_meshes;      /* the staging launch of this call (both single meshes: vertices, boxes, Morton keys) */
    float pad_ms;
    int32_t bvh_chunk_launches; /* 0 or 1 */
    uint32_t bvh_chunk_leaves;  /* leaves that launch built or refit */
    int32_t graph;              /* 1: the launches up to the counter read-back replayed a HIP graph */
    int32_t pad;
} gi_stats;

typedef struct {
    uint64_t n_pairs;
    uint32_t* pairs;        /* 2 * n_pairs: (a, b), sorted, unique */
    uint32_t* seg_offsets;  /* n_pairs + 1: segments of pair i are [off[i], off[i+1]) */
    uint64_t n_segments;
    int32_t exact_limbs;    /* 64-bit limbs per exact number (2 for grid meshes, 4 for general
                               doubles, more when host-stage pairs need them) */
    int32_t on_device;      /* arrays are device pointers */
    uint64_t* seg_exact;    /* n_segments * 8 * exact_limbs: x0 y0 z0 w0 x1 y1 z1 w1,
                               little-endian two's complement; point = (x/w, y/w, z/w)
                               * 2^seg_exp2, w > 0 */
    int32_t* seg_exp2;      /* n_segments power-of-two scales (NULL: all zero) */
    double* seg_f64;        /* n_segments * 6: endpoints rounded to double */
    gi_stats stats;
    void* internal;
} gi_result;

int gi_create(const gi_config* cfg, gi_context** out);
int gi_destroy(gi_context* ctx);
const char* gi_last_error(const gi_context* ctx);
const char* gi_status_string(int status);

/* Host arrays: xyz = 3*nv doubles, tri = 3*nt vertex indices. Copied to the
 * device before the call returns (by DMA when the arrays are pinned); the
 * index range is checked on the device, and an index >= nv leaves the slot
 * unset (GI_ERR_INVALID_ARGUMENT). */
int gi_set_mesh(gi_context* ctx, int slot, const double* xyz, uint32_t nv, const uint32_t* tri, uint32_t nt);
/* Same with device arrays (already resident in HBM). Copied device-to-device. */
int gi_set_mesh_device(gi_context* ctx, int slot, const double* d_xyz, uint32_t nv, const uint32_t* d_tri,
                       uint32_t nt);
/* Rigid motion about `center`: p' = R (p - c) + c + t with R = Rz Ry Rx from
 * Euler angles in degrees, applied as one affine map the way the reference's
 * CGAL path does (src/CPU/RotationTools.h:28-104). snap_to_grid rounds the
 * moved vertices to integers. All zero = identity. Refits the BVH. */
int gi_set_transform(gi_context* ctx, int slot, const double rot_deg[3], const double trans[3], const double center[3],
                     int snap_to_grid);
/* Same with an explicit row-major 3x4 affine matrix. */
int gi_set_transform_matrix(gi_context* ctx, int slot, const double m[12], int snap_to_grid);
/* Builds (or refits) whatever is stale. gi_intersect calls it as needed. */
int gi_build(gi_context* ctx);
/* Copies the slot's vertices as the predicates see them (after the rigid
 * transform, and the grid snap when set) to xyz (3 * nv doubles; nv must be
 * the slot's vertex count); builds first when stale. No reference
 * counterpart: it exposes the doubles the reference's CGAL path builds with
 * createRigidTransformation (src/CPU/RotationTools.h:76-104) so that callers
 * can check them. */
int gi_get_vertices(gi_context* ctx, int slot, double* xyz, uint32_t nv);
/* Copies the slot's Morton order and BVH as the traversal reads them, for
 * callers that check them; builds first when stale. n is the slot's triangle
 * count. keys, ids: the sorted Morton keys and the triangle ids in that order
 * (n each). The tree (mesh A, or a B slot that builds one): per internal node
 * i < n - 1, left[i] its left-child word (bit 31 set: the right child is a
 * leaf; low 31 bits: the left child, internal node j or leaf k as node
 * n - 1 + k; the right child is the node after the left child's range),
 * rope[i] the node after its subtree in pre-order (0xFFFFFFFF at the end) and
 * range[2 i], range[2 i + 1] its first and last leaf; leaf_rope[k] (n) the
 * same for leaf k. Any pointer may be null; a tree pointer for a slot that
 * builds no tree is GI_ERR_NOT_READY. No reference counterpart: the
 * reference's cuBQL tree (include/third-party/cubql/level_cut_builder.h) is
 * internal state; this lets tests hold the tree to Karras' definition. */
int gi_get_tree(gi_context* ctx, int slot, uint32_t n, uint32_t* keys, uint32_t* ids, uint32_t* left,
                uint32_t* rope, uint32_t* range, uint32_t* leaf_rope);
int gi_intersect(gi_context* ctx, const gi_options* opt, gi_result* out);
int gi_result_free(gi_result* res);

/* Per-call pipeline: copies both meshes, builds, intersects, frees device state. */
int gi_intersect_oneshot(const gi_config* cfg, const double* xyzA, uint32_t nvA, const uint32_t* triA, uint32_t ntA,
                         const double* xyzB, uint32_t nvB, const uint32_t* triB, uint32_t ntB,
                         const gi_options* opt, gi_result* out);

/* ------------------------------------------------------------- no allocation
 * gi_intersect_noalloc: no counterpart in the reference. It is for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same result as gi_intersect_oneshot
 * (sorted pairs and exact segments of two meshes given as host arrays),
 * computed without any allocation by the library. Every intermediate lives in
 * one static device array of GI_NOALLOC_SCRATCH_MIB MiB, which each GPU the
 * library runs on holds until the process ends; work runs on the null
 * stream; results go to the caller's host buffers.
 * Limits: at most GI_NOALLOC_MAX_TRIANGLES triangles and
 * GI_NOALLOC_MAX_VERTICES vertices per mesh (else GI_ERR_TOO_LARGE before any
 * GPU work), and candidate lists and results within what the array holds
 * after both meshes (GI_ERR_TOO_LARGE). Pairs off the integer grid that would
 * need the host residual stage are refused (GI_ERR_TOO_LARGE; use
 * gi_intersect_oneshot). When the caller's buffers are too small the call
 * returns GI_ERR_TOO_LARGE with *n_pairs / *n_segments set to what is needed.
 * Calls share each GPU's array: calls that overlap on one GPU race, so
 * callers serialise them per GPU.
 *   pairs:       2 * pairs_cap (a, b), sorted, unique
 *   seg_offsets: pairs_cap + 1 (always written: at least the leading 0)
 *   seg_exact:   seg_cap * 8 * 4 words; written as in gi_result with
 *                *exact_limbs (2 or 4) words per number
 *   seg_exp2:    seg_cap scales (zero for grid meshes)
 *   stats:       may be NULL; stage counts, no timings */
#define GI_NOALLOC_SCRATCH_MIB 128
#define GI_NOALLOC_MAX_TRIANGLES (1u << 17)
#define GI_NOALLOC_MAX_VERTICES (1u << 18)
int gi_intersect_noalloc(int32_t device, const double* xyzA, uint32_t nvA, const uint32_t* triA, uint32_t ntA,
                         const double* xyzB, uint32_t nvB, const uint32_t* triB, uint32_t ntB, int32_t want_segments,
                         uint32_t* pairs, uint64_t pairs_cap, uint32_t* seg_offsets, uint64_t* seg_exact,
                         int32_t* seg_exp2, uint64_t seg_cap, uint64_t* n_pairs, uint64_t* n_segments,
                         int32_t* exact_limbs, gi_stats* stats);

/* ------------------------------------------------------------- batches
 * Many independent mesh pairs in one pass: every stage (vertex preparation,
 * boxes, Morton sort, BVH build, traversal, filter, exact stage, result sort)
 * runs once over the whole batch instead of once per pair. Replaces the
 * reference's per-pair loop of its batch driver
 * (apps/archive/main_massive.cpp:237-304, one runGridForestIntersectionPipeline
 * per pair) and the batchMultiplier chunk loop
 * (src/single_tree/SingleTreeBatchIntersector.cu:66-73, 170-455).
 * Meshes are taken as placed (no transform); host arrays, copied. */
typedef struct {
    const double* xyzA;     /* 3 * nvA */
    const uint32_t* triA;   /* 3 * ntA, indices into this mesh's vertices */
    uint32_t nvA, ntA;
    const double* xyzB;
    const uint32_t* triB;
    uint32_t nvB, ntB;
} gi_mesh_pair;

/* Replaces both mesh slots with the batch (at most 2^30 - 1 triangles per
 * side in total and 2^29 pairs, else GI_ERR_TOO_LARGE). gi_intersect then
 * refuses the context; use: */
int gi_set_batch(gi_context* ctx, const gi_mesh_pair* pairs, uint32_t n_pairs);
/* One result for the whole batch: the pairs of mesh pair i are
 * [pair_offsets[i], pair_offsets[i+1]) (n_pairs + 1 entries, caller-owned),
 * with triangle indices local to that pair's meshes, sorted, and their
 * segments as in gi_intersect. */
int gi_intersect_batch(gi_context* ctx, const gi_options* opt, gi_result* out, uint64_t* pair_offsets);

/* Makes dst's slot a read-only view of src's (built) slot: no copy; src must
 * outlive dst's use of it. Used to serve many queries against one resident
 * BVH (no reference counterpart: the reference keeps one controller per pair).
 * A B slot builds no BVH (its triangles, in Morton order, are the queries);
 * shared into an A slot, src's B builds its tree first (GI_ERR_INVALID_ARGUMENT
 * when src's B slot is itself a view). */
int gi_share_mesh(gi_context* dst, int dst_slot, gi_context* src, int src_slot);

/* ------------------------------------------------------------- tuning
 * Launch shapes of the context's queries. No reference counterpart: the
 * reference's hand-set tunables (level-cut levels, leafThreshold,
 * batchMultiplier, dual-tree steps; apps/mainApplication.cpp:827-839) shape an
 * algorithm this library does not use. One triangle per leaf and stackless
 * rope traversal leave no leaf size or stack depth to tune. */
typedef struct {
    int32_t traverse_variant; /* 0..3: (workgroup size, candidate pairs staged per wave before one
                                 queue append) = (64,256), (128,256), (256,256), (256,128) */
    int32_t filter_blocks;    /* workgroups of the grid-stride FP64 filter launch */
    int32_t exact_blocks;     /* workgroups of the grid-stride exact-stage launch */
    int32_t build_chunk;      /* sorted leaves per BVH build workgroup: 256 or 512 */
} gi_tuning;

int gi_get_tuning(const gi_context* ctx, gi_tuning* out);
int gi_set_tuning(gi_context* ctx, const gi_tuning* tuning);
/* Times every candidate value of each knob in turn on the context's current
 * meshes (reps queries each, median of the affected stage's time) and keeps
 * the fastest; results never depend on the tuning. */
int gi_autotune(gi_context* ctx, const gi_options* opt, int32_t reps, gi_tuning* out);

/* Frees all device memory of the context but keeps the host copies; the next
 * call rebuilds (reference: clearGPU / reconstructGPU). */
int gi_clear_device(gi_context* ctx);
int gi_reconstruct_device(gi_context* ctx);

/* ------------------------------------------------------------- serving
 * A request loop around one resident static mesh (mesh A, built once): each
 * request names a registered moving mesh and a rigid transform (row-major
 * 3x4); the moving mesh is transformed and its BVH refit (same topology),
 * then intersected with the static one. `slots` requests run on the GPU at
 * once, each on its own stream; requests that arrive meanwhile queue and are
 * admitted as slots free up and while free HBM stays above the reserve.
 * No reference counterpart (the reference runs one query at a time from its
 * GUI loop, apps/mainApplication.cpp:1693-1722). */
typedef struct gi_server gi_server;

typedef struct {
    int32_t device;
    int32_t slots;              /* requests in flight on the GPU, 0 = 4 */
    int32_t snap_to_grid;       /* round moved vertices to integers (GPU exact stage) */
    int32_t want_segments;
    uint64_t hbm_reserve_bytes; /* admission: free HBM to keep, 0 = 2 GiB; a request is
                                   admitted anyway when no other request runs, so
                                   UINT64_MAX serialises the requests */
} gi_server_config;

typedef struct {
    double queue_ms;   /* submit -> admitted to a slot */
    double service_ms; /* admitted -> result ready */
    double latency_ms; /* submit -> result ready */
    double t_admit_ms; /* admission time, ms since gi_server_create */
    double t_done_ms;  /* completion time, ms since gi_server_create */
    int32_t slot;
    int32_t status;    /* gi_status of the request */
    uint32_t refusals; /* admission checks this request failed (free HBM below the reserve) */
    uint32_t pad;
} gi_request_info;

int gi_server_create(const gi_server_config* cfg, const double* xyz, uint32_t nv, const uint32_t* tri, uint32_t nt,
                     gi_server** out);
/* registers a moving mesh (uploaded once per slot); returns its id */
int gi_server_add_mesh(gi_server* srv, const double* xyz, uint32_t nv, const uint32_t* tri, uint32_t nt,
                       int32_t* mesh_id);
/* non-blocking; m = row-major 3x4 affine map applied to the registered vertices */
int gi_server_submit(gi_server* srv, int32_t mesh_id, const double m[12], uint64_t* ticket);
/* blocks until the ticket's result is ready; out holds host arrays
 * (gi_result_free), info may be NULL */
int gi_server_wait(gi_server* srv, uint64_t ticket, gi_result* out, gi_request_info* info);
int gi_server_destroy(gi_server* srv);

/* Exact intersection of one triangle pair on the host (the residual stage).
 * a, b: 9 doubles each. seg: up to 6 segments * 8 numbers * limbs_cap limbs.
 * Returns GI_OK and fills *hit, *nseg, *limbs (limbs used), *exp2. */
int gi_exact_pair_host(const double a[9], const double b[9], int32_t* hit, int32_t* coplanar, int32_t* nseg,
                       int32_t* exp2, int32_t* limbs, uint64_t* seg, int32_t limbs_cap);

#ifdef __cplusplus
}
#endif

#endif
