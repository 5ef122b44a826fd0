/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of lvr2's AdaptiveKSearchSurface hot path (uos/lvr2 25.2.4), used
 * by tests/, by the parity check in __graft_entry__.smoke() and as the timed
 * cpu_baseline in bench.py.  Nothing in the shipped library links or calls it.
 *
 * Two kNN orderings are provided:
 *   ORC_METRIC_CONTRACT  key = (fp32 d^2 computed as (dx*dx + dy*dy) + dz*dz with no
 *                        FMA contraction, original point index); the k smallest keys.
 *                        This is the ordering the GPU path is bit-exact against.
 *   ORC_METRIC_REFERENCE fp64 d^2 of (double q - double p), strict '<' insertion into a
 *                        std::priority_queue, ties left to the heap, exactly as
 *                        include/lvr2/algorithm/KDTree.tcc:186-248 does.
 */
#ifndef LVR2_ORACLE_H
#define LVR2_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_METRIC_CONTRACT = 0, ORC_METRIC_REFERENCE = 1 };

typedef struct orc_surface orc_surface;

/* AdaptiveKSearchSurface ctor (AdaptiveKSearchSurface.tcc:56-91) with the "lvr2"
 * SearchKDTree (KDTree.tcc:316-329, leaf size 20).  Copies the points. */
orc_surface* orc_surface_create(const float* xyz, int64_t n, int metric);
void orc_surface_destroy(orc_surface* s);
/* PointsetSurface ctor bounding box (PointsetSurface.tcc:52-56): min[3], max[3]. */
void orc_surface_bbox(const orc_surface* s, float* out6);

/* SearchKDTree::kSearch (KDTree.tcc:331-366): ascending; rows padded with -1 / inf
 * when fewer than k points exist. */
int orc_knn(const orc_surface* s, const float* q, int64_t m, int k,
            int32_t* out_idx, float* out_d2);

/* SearchKDTree::radiusSearch (KDTree.tcc:368-386) under the contract ordering: the k
 * smallest keys among the points with fp32 d^2 < fp32 (r * r); rows padded with -1 / inf,
 * out_count[i] = neighbours found (out_d2 / out_count may be NULL). */
int orc_radius(const orc_surface* s, const float* q, int64_t m, int k, float r, int32_t* out_idx, float* out_d2,
               int32_t* out_count);

/* PointsetSurface::getInterpolatedNormal (PointsetSurface.tcc:61-74) at m positions:
 * ((0,1,0) + sum of the ki nearest points' normalised normals) / ki.  out: m*3 floats. */
int orc_interpolated_normal(const orc_surface* s, const float* normals, int ki, const float* q, int64_t m,
                            float* out);

/* O(N*M) brute force under the contract ordering; pins orc_knn on small inputs. */
int orc_knn_brute(const float* xyz, int64_t n, const float* q, int64_t m, int k,
                  int32_t* out_idx, float* out_d2);

/* calculateSurfaceNormals without the interpolation tail
 * (AdaptiveKSearchSurface.tcc:158-284).  method: 0 LS plane (tcc:467-520),
 * 2 iterative (tcc:522-607), 3 IPCA exact / PCA (tcc:609-673).
 * subset: optional list of point ids to process (NULL = all points); outputs are
 * indexed like the subset (or like the points when subset is NULL).
 * out_n32: FP32 restatement.  out_n64 (optional): the same neighbourhoods with the
 * plane fit and the flip done in FP64.  out_k (optional): final adaptive k.
 * out_gap (optional, method 3): FP64 (lambda_mid - lambda_min) / lambda_max. */
int orc_normals(const orc_surface* s, int kn, int method, const float flip[3],
                const int64_t* subset, int64_t nsub,
                float* out_n32, double* out_n64, int32_t* out_k, double* out_gap);

/* interpolateSurfaceNormals (tcc:293-335): n_i' = normalize(n_i + sum_{j in kNN(ki)} n_j). */
int orc_interpolate(const orc_surface* s, const float* normals_in, int ki,
                    const int64_t* subset, int64_t nsub, float* out);

/* The same interpolation in fp64 (fp64 input normals, fp64 sums and normalisation). */
int orc_interpolate_f64(const orc_surface* s, const double* normals_in, int ki,
                        const int64_t* subset, int64_t nsub, double* out);

/* AdaptiveKSearchSurface::distance in fp64 on the same neighbourhoods (fp64 normals). */
int orc_distance_f64(const orc_surface* s, const double* normals, int kd,
                     const float* q, int64_t m, double* out_proj, double* out_eucl);

/* AdaptiveKSearchSurface::distance (tcc:414-457) at m query positions. */
int orc_distance(const orc_surface* s, const float* normals, int kd,
                 const float* q, int64_t m, float* out_proj, float* out_eucl);

/* PointsetGrid ctor vertex set (PointsetGrid.tcc:41-109 + HashGrid::addBox,
 * HashGrid.tcc:235-259).  Cells are added in ascending (x,y,z) index order, so the
 * vertex numbering is deterministic.  Two-call protocol: call with NULL outputs to get
 * the counts, then with buffers.  out_vtx: nv*3 floats; out_cells: nc*8 vertex ids;
 * out_cell_idx: nc*3 ints. */
int orc_grid(const float* xyz, int64_t n, const float bbox6[6], float voxel, int extrude,
             int64_t* nv, int64_t* nc, float* out_vtx, int32_t* out_cells, int32_t* out_cell_idx);

/* PointsetGrid::calcDistanceValues (PointsetGrid.tcc:111-185): invalid flags from the
 * euclidean distances, then the mask of cells that survive the prune. */
int orc_calc_distance_values(const float* eucl, int64_t nv, float voxel,
                             const int32_t* cells, int64_t nc,
                             uint8_t* out_invalid, uint8_t* out_cell_keep);

/* FastReconstruction::getMesh (FastReconstruction.tcc:47-65) on the lattice (vertices, cells from
 * orc_grid) and the signed distance of each lattice vertex, after the erase of the cells with an
 * invalid corner (invalid may be NULL): FastBox cells (FastBox.tcc:206-269) when pmc is 0,
 * BilinearFastBox cells (BilinearFastBox.tcc:56-137: faces with two equal corner positions are
 * skipped) when pmc is 1.  Cells in the given order; vertices numbered by creation.  Outputs up to
 * cap_v vertices / cap_f faces (out_face_cell, optional: the cell of each face); returns 1 when
 * either did not fit (the counts are exact either way). */
int orc_mesh(const float* vtx, int64_t nv, const int32_t* cells, int64_t nc, const float* dist,
             const uint8_t* invalid, int pmc, int64_t cap_v, int64_t cap_f, int64_t* out_nv, int64_t* out_nf,
             float* out_vtx, int32_t* out_faces, int32_t* out_face_cell);

/* BilinearFastBox::optimizePlanarFaces(mesh, kc) over all cells (BilinearFastBox.tcc:140-212), in
 * place on vtx, sequentially in face-cell order with kSearch on s.  out_moves: end-point moves
 * made; out_nonmanifold: edges with more than two faces plus half-edges used twice (0 for a mesh
 * HalfEdgeMesh accepts). */
int orc_pmc_optimize(const orc_surface* s, int kc, float* vtx, int64_t nv, const int32_t* faces,
                     const int32_t* face_cell, int64_t nf, int64_t* out_moves, int64_t* out_nonmanifold);

/* Symmetric 3x3 eigen-decomposition used for method 3 (Jacobi; FP32 and FP64),
 * exposed so the tests can pin it against numpy.linalg.eigh. cov: row-major 3x3.
 * evals[3], evecs row-major with eigenvector j in column j. */
void orc_eig3_f32(const float* cov, float* evals, float* evecs);
void orc_eig3_f64(const double* cov, double* evals, double* evecs);

/* Thread count used by the OpenMP loops (omp_get_max_threads()). */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
