// The PointsetGrid lattice on the device: the query vertices of the signed distance.
//
// Reference: PointsetGrid's constructor (PointsetGrid.tcc:41-109) marks every cell
// floor(p / voxel) (HashGrid::calcIndex, HashGrid.hpp:263-281) whose centre (i + 0.5) voxel
// (HashGrid.hpp:283-287) lies in the cloud's bounding box, optionally adds the 26 neighbours
// of each (again only those whose centre is in the box), then HashGrid::addLatticePoints /
// addBox (HashGrid.tcc:235-259, 333-378) visits the cells and creates each cell's 8 corners at
// centre + box_creation_table[v] * voxel / 2 (FastReconstructionTables.hpp:86-95), a corner
// shared with an earlier cell being reused.  The reference visits the cells in the iteration
// order of an std::unordered_set that its OpenMP threads fill under a critical section
// (PointsetGrid.tcc:59-108), so its vertex numbering, and which cell a shared corner's fp32
// position comes from, change from run to run.  This lattice fixes that order to increasing
// (x, y, z) cell index, the order the CPU restatement uses: the vertex and cell sets equal the
// reference's, the numbering and last-ulp corner positions match one of its possible orders.
//
// Here: cell keys are sorted and deduplicated with the radix sort; the 8 nc (corner, order)
// pairs are sorted by corner so that equal corners are adjacent, each corner's first use
// (the smallest order = cell rank * 8 + corner slot) gives both its vertex id (vertices are
// numbered by first use, as the sequential insertion does) and its position (computed from
// that first cell, as the reference computes it).
#pragma once

#include "common.hpp"
#include "lbvh.hpp"

namespace lvrhip {

constexpr uint64_t LAT_SENTINEL = ~0ull;
constexpr int LAT_BIAS = 1 << 20;  // cell indices must lie in (-2^20, 2^20)

// box_creation_table (FastReconstructionTables.hpp:86-95): corner v of a cell at
// centre + (bx, by, bz) * voxel / 2
__device__ __forceinline__ int lat_bct(int v, int axis)
{
    const int x = (v == 1 || v == 2 || v == 5 || v == 6) ? 1 : -1;
    const int y = (v == 2 || v == 3 || v == 6 || v == 7) ? 1 : -1;
    const int z = v >= 4 ? 1 : -1;
    return axis == 0 ? x : axis == 1 ? y : z;
}

__device__ __forceinline__ uint64_t lat_pack(int x, int y, int z)
{
    return ((uint64_t)(uint32_t)(x + LAT_BIAS) << 42) | ((uint64_t)(uint32_t)(y + LAT_BIAS) << 21) |
           (uint64_t)(uint32_t)(z + LAT_BIAS);
}
__device__ __forceinline__ void lat_unpack(uint64_t k, int& x, int& y, int& z)
{
    x = (int)(k >> 42) - LAT_BIAS;
    y = (int)((k >> 21) & ((1ull << 21) - 1)) - LAT_BIAS;
    z = (int)(k & ((1ull << 21) - 1)) - LAT_BIAS;
}

struct LatBox {
    float lo[3], hi[3];
    float voxel;
};

// BoundingBox::contains (BoundingBox.tcc:166-171) of the cell centre
__device__ __forceinline__ bool lat_contains(const LatBox& b, int x, int y, int z)
{
    const float cx = ((float)x + 0.5f) * b.voxel, cy = ((float)y + 0.5f) * b.voxel, cz = ((float)z + 0.5f) * b.voxel;
    return b.lo[0] <= cx && b.hi[0] >= cx && b.lo[1] <= cy && b.hi[1] >= cy && b.lo[2] <= cz && b.hi[2] >= cz;
}

// cell of every point (sentinel for padding and for cells whose centre is outside the box)
__global__ __launch_bounds__(256) void lat_point_cells_kernel(const float* __restrict__ px, const float* __restrict__ py,
                                                              const float* __restrict__ pz,
                                                              const uint32_t* __restrict__ pidx, int64_t n, LatBox b,
                                                              uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = LAT_SENTINEL;
    if (pidx[i] != ~0u) {
        const int x = (int)floorf(px[i] / b.voxel), y = (int)floorf(py[i] / b.voxel), z = (int)floorf(pz[i] / b.voxel);
        if (lat_contains(b, x, y, z)) k = lat_pack(x, y, z);
    }
    keys[i] = k;
    vals[i] = (uint32_t)i;
}

// 1 at the first entry of every run of equal keys (sentinels excluded)
__global__ void lat_first_flags_kernel(const uint64_t* __restrict__ keys, int64_t n, uint32_t* __restrict__ flags)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    flags[i] = (k != LAT_SENTINEL && (i == 0 || keys[i - 1] != k)) ? 1u : 0u;
}

__global__ void lat_gather_keys_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ at,
                                       const uint32_t* __restrict__ count, uint64_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (int64_t)*count) out[i] = keys[at[i]];
}

// the 27 cells around every cell (itself included), centre-in-box filtered (PointsetGrid.tcc:77-100)
__global__ __launch_bounds__(256) void lat_extrude_kernel(const uint64_t* __restrict__ cells, int64_t nc, LatBox b,
                                                          uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc * 27) return;
    const int64_t c = i / 27;
    const int d = (int)(i % 27);
    int x, y, z;
    lat_unpack(cells[c], x, y, z);
    x += d / 9 - 1;
    y += (d / 3) % 3 - 1;
    z += d % 3 - 1;
    keys[i] = lat_contains(b, x, y, z) ? lat_pack(x, y, z) : LAT_SENTINEL;
    vals[i] = (uint32_t)i;
}

// corner key of (cell rank r, slot v), value = its order r * 8 + v
__global__ __launch_bounds__(256) void lat_corners_kernel(const uint64_t* __restrict__ cells, int64_t nc,
                                                          uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc * 8) return;
    const int v = (int)(i % 8);
    int x, y, z;
    lat_unpack(cells[i / 8], x, y, z);
    keys[i] = lat_pack(x + (lat_bct(v, 0) + 1) / 2, y + (lat_bct(v, 1) + 1) / 2, z + (lat_bct(v, 2) + 1) / 2);
    vals[i] = (uint32_t)i;
}

// group (corner) number of every sorted pair: the inclusive count of run starts - 1, with the
// block offsets from the flag block sums
constexpr int LAT_THREADS = 256;
constexpr int LAT_ITEMS = 16;
constexpr int LAT_TILE = LAT_THREADS * LAT_ITEMS;

__global__ __launch_bounds__(LAT_THREADS) void lat_group_ids_kernel(const uint32_t* __restrict__ flags, int64_t n,
                                                                    const uint32_t* __restrict__ boff,
                                                                    uint32_t* __restrict__ gid)
{
    __shared__ uint32_t wc[LAT_THREADS / kWave];
    const int64_t base = (int64_t)blockIdx.x * LAT_TILE;
    const int wave = threadIdx.x / kWave;
    uint32_t run = boff[blockIdx.x];
    for (int r = 0; r < LAT_ITEMS; r++) {
        const int64_t i = base + (int64_t)r * LAT_THREADS + threadIdx.x;
        const bool f = i < n && flags[i] != 0;
        const uint64_t m = ballot(f);
        if (lane_id() == 0) wc[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t pre = 0, tot = 0;
        for (int w = 0; w < LAT_THREADS / kWave; w++) {
            if (w < wave) pre += wc[w];
            tot += wc[w];
        }
        // inclusive count of starts up to i, minus one
        const uint32_t below = (uint32_t)__popcll(m & ((2ull << lane_id()) - 1ull));
        if (i < n) gid[i] = run + pre + below - 1u;
        run += tot;
        __syncthreads();
    }
}

// first order of every corner group
__global__ void lat_first_order_kernel(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ gid,
                                       const uint32_t* __restrict__ order, int64_t n, uint64_t* __restrict__ first,
                                       uint32_t* __restrict__ gidx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flags[i]) return;
    first[gid[i]] = order[i];
    gidx[gid[i]] = gid[i];
}

// vertex id of every group: its rank in first-use order
__global__ void lat_vertex_ids_kernel(const uint32_t* __restrict__ sorted_gidx, int64_t nv, uint32_t* __restrict__ vid)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nv) vid[sorted_gidx[j]] = (uint32_t)j;
}

__global__ void lat_cells_kernel(const uint32_t* __restrict__ gid, const uint32_t* __restrict__ order,
                                 const uint32_t* __restrict__ vid, int64_t n, int32_t* __restrict__ cells)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cells[order[i]] = (int32_t)vid[gid[i]];
}

// vertex j at centre(first cell) + box_creation_table[slot] * voxel / 2 (HashGrid.tcc:333-345)
__global__ void lat_vertices_kernel(const uint64_t* __restrict__ sorted_first, const uint64_t* __restrict__ cells,
                                    int64_t nv, float voxel, float* __restrict__ vtx)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nv) return;
    const uint64_t o = sorted_first[j];
    const int v = (int)(o % 8);
    int x, y, z;
    lat_unpack(cells[o / 8], x, y, z);
    const float half = voxel / 2.0f;
    vtx[3 * j] = ((float)x + 0.5f) * voxel + (float)lat_bct(v, 0) * half;
    vtx[3 * j + 1] = ((float)y + 0.5f) * voxel + (float)lat_bct(v, 1) * half;
    vtx[3 * j + 2] = ((float)z + 0.5f) * voxel + (float)lat_bct(v, 2) * half;
}

__global__ void lat_cell_index_kernel(const uint64_t* __restrict__ cells, int64_t nc, int32_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    int x, y, z;
    lat_unpack(cells[i], x, y, z);
    out[3 * i] = x;
    out[3 * i + 1] = y;
    out[3 * i + 2] = z;
}

}  // namespace lvrhip
