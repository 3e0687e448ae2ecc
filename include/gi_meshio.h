/*
 * gi_meshio.h -- mesh input and result output of libgi.so (host code, no GPU).
 *
 *   gi_off_read         RobustIO::loadOffRobust (src/CPU/RobustOffLoader.h:116-211):
 *                       polygon soup -> fan triangulation (p0, pk, pk+1) -> rows with a
 *                       repeated vertex dropped -> later duplicates of a vertex set dropped
 *                       (the first occurrence stays) -> face i = i-th surviving triangle,
 *                       face_to_row[i] = its OFF row, vertices in file order. The
 *                       reference then builds a CGAL Surface_mesh (orienting the soup,
 *                       or dropping faces add_face rejects) and re-derives the arrays
 *                       from it; for a manifold, consistently oriented file that keeps
 *                       the face order and vertex order above. Non-manifold repairs
 *                       (vertex splitting, add_face rejection) are not reproduced: such
 *                       faces are kept as read.
 *   gi_off_write        exportMeshToOff (apps/mainApplication.cpp:464-471): OFF text,
 *                       17 significant digits.
 *   gi_pairs_csv_write  IntersectionHistory::exportRunToCsv
 *                       (src/CPU/IntersectionHistory.h:115-161): same header and columns.
 *   gi_pairs_diff_csv_write
 *                       IntersectionHistory::exportDiff (IntersectionHistory.h:163-240):
 *                       same categories, filter words and columns.
 *   gi_segments_csv_write, gi_segments_obj_write
 *                       the intersection curve (reference: IntersectionSegment doubles
 *                       handed to the viewer, src/CPU/PolyscopeBridge.h:432-466):
 *                       exact endpoints as decimal integers plus doubles, or an OBJ
 *                       polyline of the doubles.
 *
 * All functions return a gi_status; gi_meshio_last_error() describes the last
 * failure of the calling thread.
 */
#ifndef GI_MESHIO_H
#define GI_MESHIO_H

#include <stdint.h>

#include "gi_intersector.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gi_mesh_data gi_mesh_data; /* owns the arrays a gi_off_mesh points to */

typedef struct {
    uint64_t soup_points;          /* vertex rows in the file */
    uint64_t soup_polygons;        /* face rows in the file */
    uint64_t extra_fan_triangles;  /* triangles beyond the first of each polygon */
    uint64_t n_dropped_degenerate; /* entries of dropped_degenerate (a row per dropped fan triangle) */
    uint64_t n_dropped_duplicate;
    uint32_t nv, nt;
    const double* xyz;                 /* 3 * nv */
    const uint32_t* tri;               /* 3 * nt */
    const uint32_t* face_to_row;       /* nt: OFF row of face i */
    const uint32_t* dropped_degenerate;
    const uint32_t* dropped_duplicate;
} gi_off_mesh;

int gi_off_read(const char* path, gi_off_mesh* out, gi_mesh_data** owner);
int gi_off_free(gi_mesh_data* owner);
int gi_off_write(const char* path, const double* xyz, uint32_t nv, const uint32_t* tri, uint32_t nt);

/* pairs: 2 * n (faceA, faceB); coordinates are written as given (pass the moved ones) */
int gi_pairs_csv_write(const char* path, const uint32_t* pairs, uint64_t n, const double* xyzA, uint32_t nvA,
                       const uint32_t* triA, uint32_t ntA, const double* xyzB, uint32_t nvB, const uint32_t* triB,
                       uint32_t ntB);
/* filter: "all", "only1"/"1"/"exclusive1", "only2"/"2"/"exclusive2", "both" (any case) */
int gi_pairs_diff_csv_write(const char* path, const uint32_t* pairs1, uint64_t n1, uint64_t id1,
                            const uint32_t* pairs2, uint64_t n2, uint64_t id2, const char* filter, const double* xyzA,
                            uint32_t nvA, const uint32_t* triA, uint32_t ntA, const double* xyzB, uint32_t nvB,
                            const uint32_t* triB, uint32_t ntB);
/* res must hold host arrays (gi_intersect with output_on_device = 0) */
int gi_segments_csv_write(const char* path, const gi_result* res);
int gi_segments_obj_write(const char* path, const gi_result* res);

const char* gi_meshio_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
