/*
 * lvrhip — C ABI of the MI355X (gfx950) implementation of lvr2's
 * AdaptiveKSearchSurface hot path: LBVH build, exact kNN, PCA normals with adaptive
 * neighbourhoods, normal interpolation and the Hoppe signed distance at grid vertices.
 *
 * Conventions
 *  - Every function returns 0 on success, otherwise an LVRHIP_E* code;
 *    lvrhip_last_error() returns a thread-local message for the last failure.  For a
 *    device allocation failure the message contains "out of memory", the substring lvr2's
 *    LargeScaleReconstruction matches (LargeScaleReconstruction.tcc:656-662).
 *  - Host pointers are caller-owned.  Device state is owned by the handle.
 *  - Point arrays are AoS float32 xyz (n*3 floats), like lvr2's "points" channel
 *    (types/Channel.hpp:43-84).  Indices are 0-based positions in that array.
 *  - kNN order: the k smallest keys (fp32 d^2, point index), where
 *    d^2 = (dx*dx + dy*dy) + dz*dz is evaluated without FMA contraction.
 *  - All work is issued on the handle's stream (default: the null stream); functions
 *    that return host data synchronise that stream.
 */
#ifndef LVRHIP_H
#define LVRHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lvrhip_surface* lvrhip_t;

enum {
    LVRHIP_OK = 0,
    LVRHIP_EINVAL = 1,      /* bad argument */
    LVRHIP_ENOMEM = 2,      /* device allocation failed ("out of memory") */
    LVRHIP_EHIP = 3,        /* other HIP runtime error */
    LVRHIP_ESTATE = 4,      /* call order violated (e.g. distances before normals) */
    LVRHIP_ECAPACITY = 5    /* an internal fixed-size structure overflowed */
};

/* Normal estimation methods, numbered like AdaptiveKSearchSurface's calcMethod
 * (AdaptiveKSearchSurface.hpp:289-293; dispatch at AdaptiveKSearchSurface.tcc:225-246). */
enum {
    LVRHIP_METHOD_LS_PLANE = 0,   /* calcPlane, tcc:467-520 */
    LVRHIP_METHOD_ITERATIVE = 2,  /* calcPlaneIterative, tcc:522-607 */
    LVRHIP_METHOD_PCA = 3         /* calcPlaneIPCAExact, tcc:609-673 */
};

const char* lvrhip_last_error(void);

/* Replaces CudaSurface(floatArr& points, size_t num_points, int device)
 * (cuda/CudaSurface.hpp:86; CudaSurface.cu:1039-1052) and the search-tree construction of
 * AdaptiveKSearchSurface (AdaptiveKSearchSurface.tcc:76).  Copies the n points to the
 * device and builds the LBVH.  Defaults follow CudaSurface::init (CudaSurface.cu:1007-1023):
 * kn=10, ki=10, kd=5, flip point (1e6,1e6,1e6), method PCA. */
int lvrhip_create(const float* xyz, size_t n, int device, lvrhip_t* out);

/* Same, for points already resident in device memory (n*3 floats on `device`).  The build
 * reads the pointer during the call; lvrhip_rebuild and graphs captured with rebuild != 0
 * read it again, so it must stay valid for those. */
int lvrhip_create_device(const float* d_xyz, size_t n, int device, void* stream, lvrhip_t* out);

/* Batched form: many scans in one handle.  Replaces the per-scan loop of
 * estimateProjectNormals (util/ScanProjectUtils.cpp:361-410), which builds one
 * AdaptiveKSearchSurface per scan and runs calculateSurfaceNormals on it.  Scan s owns the
 * input points [scan_offsets[s], scan_offsets[s+1]) (nscans + 1 non-decreasing offsets,
 * scan_offsets[0] = 0, at most 65535 scans, empty scans allowed).  Neighbourhoods never
 * cross scans: every result equals that of a single-scan handle on the scan's points.
 * All scans share one LBVH (each scan one subtree) and each kernel launch covers every
 * scan.  Every non-empty scan must hold at least 32*kn points (and ki) when normals are
 * computed, since the reference's kSearch would otherwise return short neighbourhoods;
 * kSearch and distances are single-scan only.  The host variant keeps a device copy of
 * the points; the device variant reads d_xyz like lvrhip_create_device. */
int lvrhip_create_scans(const float* xyz, const int64_t* scan_offsets, int nscans, int device, lvrhip_t* out);
int lvrhip_create_scans_device(const float* d_xyz, const int64_t* scan_offsets, int nscans, int device, void* stream,
                               lvrhip_t* out);

/* Number of scans of the handle (1 for lvrhip_create / lvrhip_create_device). */
int lvrhip_num_scans(lvrhip_t h, int* nscans);

/* Per-scan flip points, nscans*3 floats (host): estimateProjectNormals sets each scan's
 * flip point to its scanner origin (ScanProjectUtils.cpp:392).  lvrhip_set_flip_point
 * switches back to one flip point for all scans. */
int lvrhip_set_scan_flip_points(lvrhip_t h, const float* xyz);

/* Rebuild the LBVH from the point buffer the handle reads (the device copy, or d_xyz of
 * the device variants), e.g. after the caller moved the points in place.  Invalidates
 * the normals.  No lvr2 counterpart (lvr2 builds a new surface). */
int lvrhip_rebuild(lvrhip_t h);

/* hipGraph form of the whole per-cloud pass (no lvr2 counterpart): captures, on the
 * handle's stream (or a private stream when it is the null stream), [the LBVH build from
 * the point buffer if rebuild != 0,] the adaptive normals, the interpolation (ki > 0) and
 * the copy of the normals into d_out_normals (n*3 floats, input order, device memory).
 * The capture runs the pass once eagerly first, so the graph allocates nothing; the
 * graph keeps the parameters (k, method, flip points, tuning) of its capture.
 * lvrhip_graph_launch replays it without synchronising; lvrhip_synchronize waits for it
 * and reports kNN capacity errors.  Any other computing call on the handle releases the
 * graph.  Returns LVRHIP_ESTATE while profiling is on, LVRHIP_EINVAL for ki above 64. */
int lvrhip_graph_capture(lvrhip_t h, float* d_out_normals, int rebuild);
int lvrhip_graph_launch(lvrhip_t h);
int lvrhip_synchronize(lvrhip_t h);

/* Replaces ~CudaSurface / CudaSurface::freeGPU (CudaSurface.cu:1259-1264). */
int lvrhip_destroy(lvrhip_t h);

/* Replaces CudaSurface::setKn/setKi/setKd (CudaSurface.hpp:116-142) and
 * PointsetSurface::setKn/setKi/setKd (PointsetSurface.hpp:112-126).  ki == 0 disables
 * interpolation like PointsetSurface. */
int lvrhip_set_k(lvrhip_t h, int kn, int ki, int kd);

/* Replaces CudaSurface::setFlippoint (CudaSurface.hpp:153) and
 * PointsetSurface::setFlipPoint (PointsetSurface.hpp:128-131). */
int lvrhip_set_flip_point(lvrhip_t h, float x, float y, float z);

/* Replaces CudaSurface::setMethod (CudaSurface.cu:1200-1216); takes the
 * AdaptiveKSearchSurface method number instead of a string. */
int lvrhip_set_method(lvrhip_t h, int method);

/* HIP stream for all subsequent work (hipStream_t passed as void*; NULL = null stream).
 * No counterpart in lvr2, which always uses the default stream. */
int lvrhip_set_stream(lvrhip_t h, void* stream);

/* Replaces CudaSurface::calculateNormals (CudaSurface.cu:1238-1253) with the semantics of
 * AdaptiveKSearchSurface::calculateSurfaceNormals without its interpolation tail
 * (AdaptiveKSearchSurface.tcc:158-284): adaptive k = 2kn, 4kn, ... 32kn until the
 * neighbourhood bounding box passes boundingBoxOK (tcc:337-350), the plane fit of the
 * selected method, then the flip towards the flip point. */
int lvrhip_calculate_normals(lvrhip_t h);

/* AdaptiveKSearchSurface::interpolateSurfaceNormals (tcc:293-335): each normal averaged with
 * its ki nearest points' normals.  No-op when ki == 0.  The C++ drop-in runs it inside
 * calculateNormals, where CudaSurface interpolates (CudaSurface.cu:1143-1148; its
 * interpolateNormals, CudaSurface.cu:1255-1257, is empty).  ki is clamped to the cloud's (or the
 * smallest scan's) size.  Up to 64 the kNN pass's own ki list is ranked; a larger ki searches
 * sorted kSearch rows (single-scan handles; LVRHIP_EINVAL on a multi-scan handle). */
int lvrhip_interpolate_normals(lvrhip_t h);

/* Use the given normals (n*3 floats, AoS, input order) instead of estimating them: the
 * "given normals" path of lvr2_reconstruct (src/tools/lvr2_reconstruct/Main.cpp:353,
 * 385-388), where a point buffer that already has a "normals" channel skips
 * calculateSurfaceNormals.  Enables interpolate_normals and distances. */
int lvrhip_set_normals(lvrhip_t h, const float* normals_xyz);

/* Replaces CudaSurface::getNormals(floatArr) (CudaSurface.hpp:72-80). out: n*3 floats. */
int lvrhip_get_normals(lvrhip_t h, float* out_xyz);

/* Device-resident variant: copies the normals (n*3 floats, AoS, input order) into device
 * memory `d_out` on the handle's stream without synchronising. */
int lvrhip_get_normals_device(lvrhip_t h, float* d_out);

/* Replaces SearchTree::kSearch (SearchTree.hpp:49-70; SearchKDTree::kSearch,
 * KDTree.tcc:331-348) for m queries at once.  out_idx/out_d2: m*k entries, ascending;
 * rows are padded with -1 / +inf when the cloud has fewer than k points.  out_d2 may be
 * NULL. */
int lvrhip_knn(lvrhip_t h, const float* q, size_t m, int k, int32_t* out_idx, float* out_d2);

/* Replaces SearchTree::radiusSearch (SearchTree.hpp:79-92; SearchKDTree::radiusSearch,
 * KDTree.tcc:368-386, which searches the k nearest with d^2 < r^2, KDTree.tcc:101-110, 227-237):
 * per query the k smallest keys among the points with fp32 d^2 < fp32 (r * r), ascending in
 * out_idx / out_d2 (m*k entries, padded with -1 / +inf), out_count[i] = how many were found.
 * out_d2 and out_count may be NULL.  Distances are squared (the reference returns their
 * square roots). */
int lvrhip_radius_search(lvrhip_t h, const float* q, size_t m, int k, float r, int32_t* out_idx, float* out_d2,
                         int32_t* out_count);

/* Replaces PointsetSurface::getInterpolatedNormal (PointsetSurface.tcc:61-74) for m positions:
 * (0, 1, 0) (Normal's default, Normal.hpp:62) plus the normalised normals of the ki nearest
 * points in ascending key order, divided by ki, not normalised again.  out_xyz: m*3 floats.
 * Requires normals and ki >= 1 (lvrhip_set_k). */
int lvrhip_interpolated_normals(lvrhip_t h, const float* q, size_t m, float* out_xyz);

/* Search algorithm of the searches with at most 32 neighbours (distances, kSearch, radius
 * search, interpolated normals at positions); no lvr2 counterpart.  LVRHIP_SEARCH_AUTO (the
 * default) runs them best-first, one lane per query (a batch of at most 1,024 queries one wave
 * per query); the queries whose traversal stack overflows are searched again by one wave each,
 * and any that outgrow that by the tile kernel.
 * LVRHIP_SEARCH_TILE forces the tile kernel, LVRHIP_SEARCH_WAVE the wave-per-query search for
 * every query.  All give identical results. */
enum { LVRHIP_SEARCH_AUTO = 0, LVRHIP_SEARCH_TILE = 1, LVRHIP_SEARCH_WAVE = 2 };
int lvrhip_set_search_mode(lvrhip_t h, int mode);

/* Replaces CudaSurface::distances (CudaSurface.cu:1273-1317) with the semantics of
 * AdaptiveKSearchSurface::distance (tcc:414-457) and the invalid rule of
 * PointsetGrid::calcDistanceValues (PointsetGrid.tcc:138-143): for each of the m
 * positions, proj = (v - mean_kd(p)) . normalize(mean_kd(n)), eucl = |v - mean_kd(p)|,
 * invalid = eucl > 1.75 * voxel.  Requires normals (calculate_normals first).
 * out_proj / out_eucl / out_invalid may individually be NULL. */
int lvrhip_distances(lvrhip_t h, const float* q, size_t m, float voxel, float* out_proj, float* out_eucl,
                     uint8_t* out_invalid);

/* Device-resident variant of lvrhip_distances: all pointers are device memory; runs on the
 * handle's stream without synchronising. */
int lvrhip_distances_device(lvrhip_t h, const float* d_q, size_t m, float voxel, float* d_proj, float* d_eucl,
                            uint8_t* d_invalid);

/* lvrhip_distances_device without any allocation: no counterpart in the reference; for callers
 * that cannot allocate (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as lvrhip_distances_device (bit-identical), for at
 * most LVRHIP_NOALLOC_MAX_QUERIES queries per call and kd <= 32 (LVRHIP_EINVAL otherwise).  Its
 * scratch is one static device array of LVRHIP_NOALLOC_SCRATCH_BYTES that every device holds
 * for the life of the process, shared by all handles on that device: calls that overlap on one
 * device (other streams, other threads) race and must be serialised by the caller.  Queries are
 * searched in the caller's order (spatially coherent orders, such as the lattice of
 * lvrhip_grid_build, run fastest).  Synchronises the handle's stream; a query whose search
 * outgrows the stacks (never seen in practice) is reported as LVRHIP_ECAPACITY. */
#define LVRHIP_NOALLOC_MAX_QUERIES 1048576
#define LVRHIP_NOALLOC_SCRATCH_BYTES 25167872
int lvrhip_distances_device_noalloc(lvrhip_t h, const float* d_q, size_t m, float voxel, float* d_proj,
                                    float* d_eucl, uint8_t* d_invalid);

/* Differentiation form (no lvr2 counterpart; the reverse mode of AdaptiveKSearchSurface::distance,
 * tcc:414-457, through calcPlaneIPCAExact, tcc:609-673): for the m positions d_q (device) and
 * upstream weights d_grad_proj[v] = dL/dproj_v (device), adds dL/dpoints to d_grad_xyz (n*3
 * floats, input order, device memory, accumulated into).  proj is the value lvrhip_distances
 * returns for the same positions.  Neighbour sets, adaptive k and flip signs are held fixed
 * (they are piecewise constant in the positions); the neighbour lists are recomputed, not
 * stored.  Requires normals from lvrhip_calculate_normals with the PCA method and without
 * interpolation (ki = 0), on a single-scan handle.  Runs on the handle's stream and
 * synchronises it. */
int lvrhip_distances_backward_device(lvrhip_t h, const float* d_q, size_t m, const float* d_grad_proj,
                                     float* d_grad_xyz);

/* Statistics of the last calculate_normals call: n_round[r] = number of points whose
 * neighbourhood search ran with k = kn * 2^(r+1), r = 0..4.  No lvr2 counterpart. */
int lvrhip_normals_stats(lvrhip_t h, int64_t n_round[5]);

/* The neighbourhood size each point's normal was fitted on by the last calculate_normals:
 * the k of its adaptive round (kn * 2^r, r = 1..5, at most the scan's size), the k at which
 * calculateSurfaceNormals' loop stopped (AdaptiveKSearchSurface.tcc:187-216).  n int32, input
 * order.  No lvr2 counterpart (the loop's k is local there). */
int lvrhip_get_adaptive_k(lvrhip_t h, int32_t* out_k);

/* The PointsetGrid lattice on the device (PointsetGrid.tcc:41-109 with HashGrid::addLatticePoints
 * and addBox, HashGrid.tcc:235-259, 333-378): cells floor(p / voxel) whose centre lies in the
 * cloud's bounding box, with extrude != 0 also their 26 neighbours (centre in the box); each
 * cell's 8 corners at centre + box_creation_table[v] * voxel / 2 (FastReconstructionTables.hpp:
 * 86-95), a corner shared with an earlier cell reused, vertex ids in order of first use.  The
 * reference inserts the cells in the iteration order of an unordered_set filled by OpenMP threads
 * under a critical section (PointsetGrid.tcc:59-108), which varies between runs; here the cells
 * are taken in increasing (x, y, z) index order, one of the reference's possible orders (same
 * vertex and cell sets).  Returns the vertex and cell counts; the lattice stays in the handle
 * (single-scan handles only).  Cell indices must lie within +-2^20. */
int lvrhip_grid_build(lvrhip_t h, float voxel, int extrude, size_t* nv, size_t* nc);
/* The same lattice over a given bounding box (box: min x, y, z, max x, y, z) instead of the
 * cloud's: PointsetGrid's bb argument (PointsetGrid.tcc:41-46), which LargeScaleReconstruction
 * sets to the chunk's box plus its overlap (LargeScaleReconstruction.tcc:207, 633).  Cells still
 * come from the cloud's points; a cell is kept when its centre lies in box. */
int lvrhip_grid_build_box(lvrhip_t h, float voxel, int extrude, const float* box, size_t* nv, size_t* nc);
/* Copies of the last lattice: vertices (nv*3 floats, the positions lvrhip_distances takes),
 * cells (nc*8 vertex ids per cell, corner order of box_creation_table), cell index (nc*3 int32).
 * Any output may be NULL.  _device: device pointers, on the handle's stream, not synchronised. */
int lvrhip_grid_get(lvrhip_t h, float* vertices, int32_t* cells, int32_t* cell_index);
int lvrhip_grid_get_device(lvrhip_t h, float* d_vertices, int32_t* d_cells, int32_t* d_cell_index);

/* Mesh decompositions of lvrhip_mesh_build (lvr2_reconstruct's --decomposition, Options.cpp:74). */
#define LVRHIP_MESH_MC 0  /* "MC": FastBox cells */
#define LVRHIP_MESH_PMC 1 /* "PMC", lvr2_reconstruct's default: BilinearFastBox cells */

/* Marching cubes over the lattice of the last lvrhip_grid_build: FastReconstruction::getMesh
 * (FastReconstruction.tcc:47-65, 155-165) after the erase of every cell with an invalid corner
 * that PointsetGrid::calcDistanceValues performs (PointsetGrid.tcc:153-184); replaces the
 * reconstruction->getMesh(mesh) call of lvr2_reconstruct (Main.cpp:424-452).  d_dist: the signed
 * distance of each lattice vertex (nv floats, device memory: the proj of lvrhip_distances_device
 * at the lattice's vertices); d_invalid: their invalid flags (nv bytes, device memory) or NULL
 * (no cell erased); nv must equal the lattice's vertex count.  A cell's case is its corners with distance > 0 (FastBox.tcc:107-117), its
 * triangles lvr2's marching-cubes table (MCTable.hpp); one mesh vertex per lattice edge with a
 * sign change, at FastBox::calcIntersection's position as computed by the first cell (in
 * increasing cell index order) that uses it.
 * decomposition LVRHIP_MESH_MC: FastBox::getSurface (FastBox.tcc:206-269).
 * decomposition LVRHIP_MESH_PMC: BilinearFastBox::getSurface (BilinearFastBox.tcc:56-137: a face
 * with two equal corner positions is skipped), then BilinearFastBox::optimizePlanarFaces(mesh, 5)
 * on every cell (BilinearFastBox.tcc:140-212): the end points of the border edges of a cell with
 * one or two of them move to the centroid of their 5 nearest cloud points, once per such
 * (cell, edge); the handle's cloud is the search tree.
 * The reference visits its cells in the order of an unordered map, so its vertex and face
 * numbering vary between runs; this one is fixed (vertices in order of first use, faces in cell
 * order).  Returns the vertex and face counts; the mesh stays in the handle.  Single-scan
 * handles only. */
int lvrhip_mesh_build(lvrhip_t h, const float* d_dist, const uint8_t* d_invalid, size_t nv, int decomposition,
                      size_t* nverts, size_t* nfaces);
/* lvrhip_mesh_build with the distances and invalid flags in host memory (nv floats / nv bytes,
 * invalid may be NULL). */
int lvrhip_mesh_build_host(lvrhip_t h, const float* dist, const uint8_t* invalid, size_t nv, int decomposition,
                           size_t* nverts, size_t* nfaces);
/* Copies of the last mesh: vertices (nverts*3 floats) and faces (nfaces*3 vertex ids, the
 * winding of the table).  Either may be NULL.  _device: device pointers, not synchronised. */
int lvrhip_mesh_get(lvrhip_t h, float* vertices, int32_t* faces);
int lvrhip_mesh_get_device(lvrhip_t h, float* d_vertices, int32_t* d_faces);

/* Number of points of the handle. */
int lvrhip_size(lvrhip_t h, size_t* n);

/* Device-time accounting, no lvr2 counterpart.  When enabled, the library brackets its
 * kernels with hipEvents on the handle's stream and accumulates, per timer, the number of
 * launches and the summed elapsed milliseconds (read after the call that launched them
 * returns).  Timers: */
enum {
    LVRHIP_T_NORMALS_FIRST = 0, /* kNN tile kernel of the normals, first adaptive round */
    LVRHIP_T_NORMALS_RETRY = 1, /* the same kernel in the later adaptive rounds */
    LVRHIP_T_INTERPOLATE = 2,   /* interpolation kernel */
    LVRHIP_T_SDF_KERNEL = 3,    /* fused kNN + signed distance kernel (best-first, or the tile kernel) */
    LVRHIP_T_SDF_SORT = 4,      /* Morton sort of the query vertices */
    LVRHIP_T_BUILD = 5,         /* LBVH build */
    LVRHIP_T_SDF_REDO = 6,      /* tile kernel for the vertices whose best-first search overflowed */
    LVRHIP_T_COUNT = 7
};
int lvrhip_set_profiling(lvrhip_t h, int on);

/* Kernel work counters collected while profiling is on (no lvr2 counterpart):
 * out[g * 16 + c] for group g (0..4 the normals' adaptive rounds 1..5, 5 the merged later
 * rounds and the wave-per-query pass of the signed distances, 6 SDF, 7 other searches:
 * kSearch, the ki list, the adjoint) and counter c (0 tiles,
 * 1 radius attempts, 2 LDS-resident attempts, 3 candidate points scanned in the histogram
 * pass, 4 refinement passes, 5 chunk loads, 6 queries resolved, 7 traversal node pops;
 * 8..15 shader-clock cycles summed over waves in: traversal, chunk staging, per-query leaf
 * lists, histogram scan, selection, collection scan, finish, total). */
int lvrhip_get_kernel_stats(lvrhip_t h, int64_t out[128], int reset);

/* Launch shape of the kNN tile kernels (no lvr2 counterpart): lanes per query (2, 4 or 8;
 * queries per wavefront = 64 / lanes), consecutive tiles per wavefront in the first adaptive
 * round (default 12; the later rounds take one), and the radius margin carried from one tile
 * to the next.  Values <= 0 keep the current setting. */
int lvrhip_set_tuning(lvrhip_t h, int lanes_per_query, int tiles_per_wave, float r2_margin);
/* ms[i], launches[i] for i < LVRHIP_T_COUNT; either may be NULL.  reset != 0 clears them. */
int lvrhip_get_timings(lvrhip_t h, double* ms, int64_t* launches, int reset);

/* Autotuner (no lvr2 counterpart): times the adaptive normals pass on a sample of the
 * cloud (every 16th run of 4096 points in Morton order) for lanes per query in {4, 8},
 * tiles per wavefront in {4, 8, 12, 16} and radius margins in {1.15, 1.3, 1.6}, keeps the
 * fastest launch shape and returns it (any output may be NULL), then keeps the faster of the
 * two retry modes (lvrhip_set_retry_mode) at that shape.  Invalidates the normals. */
int lvrhip_autotune(lvrhip_t h, int* lanes_per_query, int* tiles_per_wave, float* r2_margin);

/* Launch shape of the signed-distance searches (distances and its reverse pass): consecutive
 * tiles per wavefront (default 2) and radius margin (default 1.3); lanes per query are shared
 * with lvrhip_set_tuning.  Values <= 0 keep the current setting.  No lvr2 counterpart. */
int lvrhip_set_sdf_tuning(lvrhip_t h, int tiles_per_wave, float r2_margin);

/* How calculate_normals runs the adaptive rounds 2..5 for the points round 1 rejected (no lvr2
 * counterpart; the results are the same either way): merged != 0 searches k = 32 kn once with
 * the four round sizes as targets, takes the first round whose neighbourhood passes
 * boundingBoxOK and searches the points that stop before round 5 again at their round's k;
 * merged == 0 (the default) searches round by round, each round only for the points the
 * previous one rejected.  lvrhip_autotune also chooses this. */
int lvrhip_set_retry_mode(lvrhip_t h, int merged);
int lvrhip_get_retry_mode(lvrhip_t h, int* merged);

/* The current launch shape (any output may be NULL). */
int lvrhip_get_tuning(lvrhip_t h, int* lanes_per_query, int* tiles_per_wave, float* r2_margin);

#ifdef __cplusplus
}
#endif
#endif
