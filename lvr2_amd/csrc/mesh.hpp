// Marching cubes on the signed distances at the lattice vertices: FastReconstruction::getMesh
// with FastBox cells (FastReconstruction.tcc:47-65, FastBox.tcc:96-263), after
// PointsetGrid::calcDistanceValues has erased every cell with an invalid corner
// (PointsetGrid.tcc:153-184).
//
// A valid cell's case is the bit set of its corners with distance > 0 (FastBox::getIndex,
// FastBox.tcc:107-117); its triangles come from lvr2's marching-cubes table (MCTable.hpp, a
// variant of Bourke's table in lvr2's edge numbering with at most four triangles per case, packed
// in mc_table.h).
// A mesh vertex sits on a lattice edge with a sign change; lvr2 creates it in the first cell
// that uses the edge and hands it to the cells around the edge (FastBox.tcc:236-253).  Here the
// cells are taken in increasing index order (the lattice's order): the triangle corners are
// keyed by their lattice edge (low vertex id * 3 + axis), sorted stably, and the first use of
// each key creates its vertex, with the position the creating cell computes
// (FastBox::calcIntersection / getIntersections, FastBox.tcc:140-202): along the edge
// x2 - d2 (x1 - x2) / (d1 - d2) from the edge's low corner x1 to its high corner x2, the other two
// coordinates copied from the corner lvr2's getIntersections names for that edge.  Vertices are
// numbered in order of first use and faces in cell order, as a sequential loop creates them.
//
// Planar marching cubes (PMC, lvr2_reconstruct's default decomposition: BilinearFastBox cells,
// BilinearFastBox.tcc:56-137) differs in two steps.  A face two of whose corner positions, as
// the cell computes them, are equal is skipped before any of its vertices is created
// (BilinearFastBox.tcc:83-97).  After all faces exist, every cell whose own faces have one or
// two border edges (an edge with a face on one side only) moves both end points of each such
// edge to the centroid of the 5 cloud points nearest to the end point's current position
// (optimizePlanarFaces, BilinearFastBox.tcc:140-212, called from FastReconstruction.tcc:155-165).
// A move reads nothing but the moved vertex's own position, so the sequential loop's result is
// the same in any cell order: a vertex named by c such (cell, edge) pairs moves c times.  Here
// round r moves every vertex with c >= r once, with one k = 5 search per vertex.
#pragma once

#include "common.hpp"
#include "mc_table.h"

namespace lvrhip {

__constant__ uint64_t MC_TRI[256] = {LVRHIP_MC_TABLE};

// cell-local edge e: corners (lo, hi) with lo below hi on the axis, and the corner whose other two
// coordinates the vertex takes (FastBox.tcc:159-202)
__constant__ int8_t MC_EDGE[12][4] = {
    {0, 1, 0, 0}, {1, 2, 1, 1}, {3, 2, 0, 2}, {0, 3, 1, 3}, {4, 5, 0, 4}, {5, 6, 1, 5},
    {7, 6, 0, 6}, {4, 7, 1, 7}, {0, 4, 2, 0}, {1, 5, 2, 1}, {3, 7, 2, 3}, {2, 6, 2, 2},
};

__device__ __forceinline__ int mc_ntri(uint64_t w) { return (int)(w >> 60); }
__device__ __forceinline__ int mc_edge(uint64_t w, int k) { return (int)((w >> (4 * k)) & 0xF); }

// FastBox::calcIntersection (FastBox.tcc:140-160), evaluated as written in fp32
__device__ __forceinline__ float mc_intersection(float x1, float x2, float d1, float d2)
{
    if (d1 == 0.0f) return x1;
    if (d2 == 0.0f) return x2;
    if ((d1 < 0.0f && d2 > 0.0f) || (d2 < 0.0f && d1 > 0.0f)) return x2 - d2 * (x1 - x2) / (d1 - d2);
    return 0.0f;
}

// the position cell cv computes for its edge e (FastBox::getIntersections, FastBox.tcc:159-202)
__device__ __forceinline__ float3 mc_edge_point(const int32_t* __restrict__ cv, int e, const float* __restrict__ vtx,
                                                const float* __restrict__ dist)
{
    const int64_t a = cv[MC_EDGE[e][0]], b = cv[MC_EDGE[e][1]], s = cv[MC_EDGE[e][3]];
    const int ax = MC_EDGE[e][2];
    const float x = mc_intersection(vtx[3 * a + ax], vtx[3 * b + ax], dist[a], dist[b]);
    return make_float3(ax == 0 ? x : vtx[3 * s], ax == 1 ? x : vtx[3 * s + 1], ax == 2 ? x : vtx[3 * s + 2]);
}

__device__ __forceinline__ bool mc_same(float3 a, float3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// per cell: its case (0 when a corner is invalid: no triangles), the mask of the table's
// triangles it keeps (all of them, or with pmc the non-degenerate ones) and their count
__global__ void mc_cells_kernel(const int32_t* __restrict__ cells, int64_t nc, const float* __restrict__ vtx,
                                const float* __restrict__ dist, const uint8_t* __restrict__ invalid, int pmc,
                                uint8_t* __restrict__ mcase, uint8_t* __restrict__ keep, uint32_t* __restrict__ ntri)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int32_t* cv = cells + 8 * c;
    int idx = 0;
    bool bad = false;
#pragma unroll
    for (int v = 0; v < 8; v++) {
        const int32_t id = cv[v];
        bad = bad || (invalid && invalid[id] != 0);
        if (dist[id] > 0.0f) idx |= 1 << v;
    }
    if (bad) idx = 0;
    const uint64_t w = MC_TRI[idx];
    const int nt = mc_ntri(w);
    uint32_t k = (1u << nt) - 1u;
    if (pmc) {
        for (int t = 0; t < nt; t++) {
            const float3 p0 = mc_edge_point(cv, mc_edge(w, 3 * t), vtx, dist);
            const float3 p1 = mc_edge_point(cv, mc_edge(w, 3 * t + 1), vtx, dist);
            const float3 p2 = mc_edge_point(cv, mc_edge(w, 3 * t + 2), vtx, dist);
            if (mc_same(p0, p1) || mc_same(p0, p2) || mc_same(p2, p1)) k &= ~(1u << t);
        }
    }
    mcase[c] = (uint8_t)idx;
    keep[c] = (uint8_t)k;
    ntri[c] = (uint32_t)__popc(k);
}

// total = the last exclusive offset + the last count
__global__ void mc_total_kernel(const uint32_t* __restrict__ off, const uint32_t* __restrict__ ntri, int64_t nc,
                                uint32_t* __restrict__ total)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *total = nc > 0 ? off[nc - 1] + ntri[nc - 1] : 0u;
}

// the three corners of every kept triangle: key = lattice edge, value = order f * 3 + j; and the
// cell of every face
__global__ void mc_emit_kernel(const int32_t* __restrict__ cells, int64_t nc, const uint8_t* __restrict__ mcase,
                               const uint8_t* __restrict__ keep, const uint32_t* __restrict__ off,
                               uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                               uint32_t* __restrict__ face_cell)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const uint64_t w = MC_TRI[mcase[c]];
    const int nt = mc_ntri(w);
    const uint32_t k = keep[c];
    uint32_t f = off[c];
    for (int t = 0; t < nt; t++) {
        if (!((k >> t) & 1u)) continue;
        face_cell[f] = (uint32_t)c;
        for (int j = 0; j < 3; j++) {
            const int e = mc_edge(w, 3 * t + j);
            const uint64_t lo = (uint64_t)(uint32_t)cells[8 * c + MC_EDGE[e][0]];
            const uint64_t o = 3 * (uint64_t)f + j;
            keys[o] = lo * 3 + (uint64_t)MC_EDGE[e][2];
            vals[o] = (uint32_t)o;
        }
        f++;
    }
}

// position of vertex j, created by corner (o % 3) of face o / 3
__global__ void mc_vertices_kernel(const uint64_t* __restrict__ first, int64_t nverts,
                                   const uint32_t* __restrict__ face_cell, const uint32_t* __restrict__ off,
                                   const uint8_t* __restrict__ mcase, const uint8_t* __restrict__ keep,
                                   const int32_t* __restrict__ cells, const float* __restrict__ vtx,
                                   const float* __restrict__ dist, float* __restrict__ out)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nverts) return;
    const uint64_t o = first[j];
    const uint64_t f = o / 3;
    const uint32_t c = face_cell[f];
    // the table triangle of the face: the (f - off[c])-th kept one
    uint32_t k = keep[c];
    for (uint32_t r = (uint32_t)(f - off[c]); r > 0; r--) k &= k - 1u;
    const int t = __ffs((int)k) - 1;
    const int e = mc_edge(MC_TRI[mcase[c]], 3 * t + (int)(o % 3));
    const float3 p = mc_edge_point(cells + 8 * (int64_t)c, e, vtx, dist);
    out[3 * j] = p.x;
    out[3 * j + 1] = p.y;
    out[3 * j + 2] = p.z;
}

// ---- planar marching cubes: the border-edge moves (BilinearFastBox::optimizePlanarFaces) -----

// the undirected edges of every face: key (low vertex id, high vertex id), value face * 3 + slot,
// slot b joining corners b and (b + 1) % 3
__global__ void mc_face_edges_kernel(const int32_t* __restrict__ faces, int64_t nf, uint64_t* __restrict__ keys,
                                     uint32_t* __restrict__ vals)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * nf) return;
    const int64_t f = i / 3;
    const int b = (int)(i % 3);
    const uint32_t u = (uint32_t)faces[3 * f + b], v = (uint32_t)faces[3 * f + (b + 1) % 3];
    keys[i] = u < v ? ((uint64_t)u << 32) | v : ((uint64_t)v << 32) | u;
    vals[i] = (uint32_t)i;
}

// after the sort by edge: a slot is a border edge when its edge has no other face
// (HalfEdgeMesh::isBorderEdge, HalfEdgeMesh.tcc:732-746)
__global__ void mc_border_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, int64_t m,
                                 uint8_t* __restrict__ border)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint64_t k = keys[i];
    const bool single = (i == 0 || keys[i - 1] != k) && (i == m - 1 || keys[i + 1] != k);
    border[vals[i]] = single ? 1 : 0;
}

// cells whose own faces have one or two border edges: one move for each end point of each; the
// largest count reached goes to *max_moves (one atomic per wave)
__global__ void mc_moves_kernel(const uint32_t* __restrict__ off, const uint32_t* __restrict__ ntri, int64_t nc,
                                const uint8_t* __restrict__ border, const int32_t* __restrict__ faces,
                                uint32_t* __restrict__ moves, uint32_t* __restrict__ max_moves)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t most = 0;
    if (c < nc) {
        const int64_t s0 = 3 * (int64_t)off[c], s1 = s0 + 3 * (int64_t)ntri[c];
        int cnt = 0;
        for (int64_t s = s0; s < s1; s++) cnt += border[s];
        if (cnt == 1 || cnt == 2) {
            for (int64_t s = s0; s < s1; s++) {
                if (!border[s]) continue;
                const int64_t f = s / 3;
                const int b = (int)(s % 3);
                const uint32_t u = (uint32_t)faces[3 * f + b], v = (uint32_t)faces[3 * f + (b + 1) % 3];
                most = max(most, atomicAdd(&moves[u], 1u) + 1u);
                most = max(most, atomicAdd(&moves[v], 1u) + 1u);
            }
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) most = max(most, (uint32_t)shfl_xor((int)most, m));
    if (lane_id() == 0 && most > 0) atomicMax(max_moves, most);
}

// round r: the vertices that move at least r times, with their current positions (any order:
// every move reads its own vertex only)
__global__ void mc_round_list_kernel(const uint32_t* __restrict__ moves, int64_t nv, uint32_t r,
                                     const float* __restrict__ vtx, uint32_t* __restrict__ list,
                                     float* __restrict__ q, uint32_t* __restrict__ count)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nv || moves[j] < r) return;
    const uint32_t o = atomicAdd(count, 1u);
    list[o] = (uint32_t)j;
    q[3 * (size_t)o] = vtx[3 * j];
    q[3 * (size_t)o + 1] = vtx[3 * j + 1];
    q[3 * (size_t)o + 2] = vtx[3 * j + 2];
}

__global__ void mc_round_scatter_kernel(const uint32_t* __restrict__ list, int64_t m, const float* __restrict__ c,
                                        float* __restrict__ vtx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const size_t j = list[i];
    vtx[3 * j] = c[3 * i];
    vtx[3 * j + 1] = c[3 * i + 1];
    vtx[3 * j + 2] = c[3 * i + 2];
}

}  // namespace lvrhip
