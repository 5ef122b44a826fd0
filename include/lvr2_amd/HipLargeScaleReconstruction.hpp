// lvr2::HipLargeScaleReconstruction: LargeScaleReconstruction's chunk loop with the GPU surface
// (LargeScaleReconstruction.tcc:85-245 and createChunk, :571-707), its out-of-memory shrink and
// retry (:649-677) and the TSDF overlap merge (:248-370), over the C ABI of liblvrhip.so.  Every
// chunk runs normals, lattice, signed distances and marching cubes on the device.
//
// What it does, step by step:
//   1. Chunks are the cells floor(p / bgVoxelSize) that hold points (BigGrid::calcIndex,
//      BigGrid.hpp:182-187), taken in increasing (x, y, z) order.
//   2. A chunk's grid box is its cell widened by overlap = 5 * voxelSize on every side
//      (tcc:145-147, 207).  Its points are the cloud's points inside that box, bounds included
//      (BigGrid::points, BigGrid.tcc:661-710), in input order.  A chunk with fewer than
//      2 * max(kn, ki, kd) points is skipped (tcc:95, 580-589).
//   3. While a chunk holds more than maxPointsPerChunk points (130,000,000 to start, tcc:98), it
//      is thinned by random sampling on a voxel grid: target voxel voxelSize / 4, doubled each
//      pass (tcc:611-620).  The reference's RandomSampleOctreeReduction seeds from
//      std::random_device, so no fixed result exists; here a fixed seed.
//   4. On the device: normals (estimate + interpolation with ki, flipped towards flipPoint, as
//      GpuSurface::calculateNormals, tcc:640-675), the PointsetGrid lattice over the grid box
//      (tcc:633), the signed distance and invalid flag of every lattice vertex, and the erase of
//      every cell with an invalid corner (PointsetGrid::calcDistanceValues, PointsetGrid.tcc:
//      111-185).  A device allocation failure ("out of memory") multiplies maxPointsPerChunk by
//      0.8 and runs the chunk again (tcc:656-671).  When the limit falls below the chunk
//      minimum, the reference switches to its CPU surface; this build has no CPU path and
//      throws instead.
//   5. With mergeChunkBorders (the default, tcc:246-370), every lattice vertex of a chunk's
//      remaining cells that another chunk's remaining cells also hold gets the weighted mean of
//      all those chunks' distances.  A chunk's weight is min(1, min over the axes of
//      (bgVoxelSize / 2 + overlap - |centre - p|) / (2 * overlap)), floored at 0.001, where
//      centre is the chunk's grid-box centroid and p the vertex's position in the first holder.
//      Sums are in double, as in the reference.  The
//      reference finds "the same vertex" by the lattice index of p + voxel / 2; here the integer
//      lattice index itself.  It sums its neighbours in OpenMP order and itself last; here every
//      holder is summed in chunk-coordinate order, with one position for the vertex, so all
//      chunks sharing a vertex get the same bits (the reference's own aim, tcc:264-266).  The
//      reference weighs with each chunk's own copy of the position, equal to this one up to the
//      last bit.
//   6. Marching cubes with FastBox cells (LargeScaleReconstruction's BoxT,
//      LargeScaleReconstruction.hpp:308) over the chunk's remaining cells, with the merged
//      distances (processChunk, tcc:712-720).  The mesh post-processing after getMesh is not
//      done here.
//
// Deviations that stay: the chunk order and the point order inside a chunk are fixed here.  The
// reference takes both from unordered-map and OpenMP order, so they vary between its runs.
#ifndef LVR2_AMD_HIPLARGESCALERECONSTRUCTION_HPP
#define LVR2_AMD_HIPLARGESCALERECONSTRUCTION_HPP

#include "../lvrhip.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <map>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace lvr2
{

// LSROptions (LargeScaleReconstruction.hpp:100-185): the fields the GPU chunk loop reads, with
// the reference's defaults; voxelSize is voxelSizes[0].
struct HipLSROptions {
    float bgVoxelSize = 10.0f;
    float voxelSize = 0.1f;
    int kn = 20, ki = 20, kd = 20;
    std::array<float, 3> flipPoint{10000000.0f, 10000000.0f, 10000000.0f};
    bool extrude = true;
    bool mergeChunkBorders = true;
    size_t maxPointsPerChunk = 130000000;
    int device = 0;
    uint64_t reductionSeed = 1;
};

struct HipLSRChunk {
    std::array<int32_t, 3> coord{};
    std::array<float, 6> gridBox{};  // min xyz, max xyz
    int retries = 0;                 // out-of-memory retries of this chunk
    size_t pointsBeforeReduction = 0;
    std::vector<float> points;       // the chunk's points after any reduction (xyz)
    std::vector<float> normals;
    std::vector<float> vertices;     // the lattice over the grid box (xyz per vertex)
    std::vector<int32_t> cells;      // 8 vertex ids per cell
    std::vector<int32_t> cellIndex;  // 3 per cell
    std::vector<float> rawDistance;  // signed distance per vertex, before the merge
    std::vector<float> distance;     // after the merge (= rawDistance without it)
    std::vector<uint8_t> invalid;    // per vertex; a cell with an invalid corner is erased
    std::vector<float> meshVertices;
    std::vector<int32_t> meshFaces;
};

namespace hip_lsr_detail
{

inline void check(int rc)
{
    if (rc != LVRHIP_OK) throw std::runtime_error(std::string("lvrhip: ") + lvrhip_last_error());
}

inline uint64_t pack3(int64_t x, int64_t y, int64_t z)
{
    const int64_t B = 1 << 20;
    return ((uint64_t)(x + B) << 42) | ((uint64_t)(y + B) << 21) | (uint64_t)(z + B);
}

// One handle with its own cleanup.
struct Handle {
    lvrhip_t h = nullptr;
    Handle(const float* xyz, size_t n, int device) { check(lvrhip_create(xyz, n, device, &h)); }
    ~Handle()
    {
        if (h) lvrhip_destroy(h);
    }
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
};

// Random sampling on a voxel grid: at most keep points per voxel of edge `voxel`, chosen by a
// seeded shuffle; the survivors keep their input order.
inline std::vector<float> reduce(const std::vector<float>& pts, float voxel, size_t keep, uint64_t seed)
{
    const size_t n = pts.size() / 3;
    std::vector<std::pair<uint64_t, uint32_t>> key(n);
    std::mt19937_64 rng(seed);
    for (size_t i = 0; i < n; i++) {
        const int64_t x = (int64_t)std::floor(pts[3 * i] / voxel), y = (int64_t)std::floor(pts[3 * i + 1] / voxel),
                      z = (int64_t)std::floor(pts[3 * i + 2] / voxel);
        key[i] = {pack3(x, y, z), (uint32_t)i};
    }
    std::sort(key.begin(), key.end());
    std::vector<char> kept(n, 0);
    for (size_t a = 0; a < n;) {
        size_t b = a;
        while (b < n && key[b].first == key[a].first) b++;
        std::vector<uint32_t> ids;
        for (size_t j = a; j < b; j++) ids.push_back(key[j].second);
        std::shuffle(ids.begin(), ids.end(), rng);
        for (size_t j = 0; j < std::min(keep, ids.size()); j++) kept[ids[j]] = 1;
        a = b;
    }
    std::vector<float> out;
    for (size_t i = 0; i < n; i++)
        if (kept[i]) out.insert(out.end(), pts.begin() + 3 * i, pts.begin() + 3 * i + 3);
    return out;
}

inline bool is_oom(const std::runtime_error& e)
{
    return std::string(e.what()).find("out of memory") != std::string::npos;
}

}  // namespace hip_lsr_detail

class HipLargeScaleReconstruction
{
public:
    explicit HipLargeScaleReconstruction(const HipLSROptions& o) : m_opt(o), m_maxPoints(o.maxPointsPerChunk) {}

    // The chunks of an n-point cloud (xyz, host memory), each with its lattice, distances and mesh.
    std::vector<HipLSRChunk> reconstruct(const float* xyz, size_t n)
    {
        using namespace hip_lsr_detail;
        m_withoutCells.clear();
        const float cs = m_opt.bgVoxelSize, voxel = m_opt.voxelSize;
        if (!(cs > 0.0f) || !(voxel > 0.0f)) throw std::runtime_error("HipLargeScaleReconstruction: sizes must be > 0");
        const size_t minPoints = 2 * (size_t)std::max(m_opt.ki, std::max(m_opt.kd, m_opt.kn));
        const float overlap = 5 * voxel;
        // 1. chunk cells and their points
        std::map<std::array<int32_t, 3>, std::vector<uint32_t>> bucket;
        for (size_t i = 0; i < n; i++) {
            const std::array<int32_t, 3> c{(int32_t)std::floor(xyz[3 * i] / cs), (int32_t)std::floor(xyz[3 * i + 1] / cs),
                                           (int32_t)std::floor(xyz[3 * i + 2] / cs)};
            bucket[c].push_back((uint32_t)i);
        }
        std::vector<HipLSRChunk> chunks;
        for (auto& kv : bucket) {
            HipLSRChunk ch;
            ch.coord = kv.first;
            // 2. the grid box and the points inside it
            float lo[3], hi[3];
            for (int a = 0; a < 3; a++) {
                const float pos = (float)ch.coord[a] * cs;
                lo[a] = pos - overlap;
                hi[a] = (pos + cs) + overlap;
                ch.gridBox[a] = lo[a];
                ch.gridBox[3 + a] = hi[a];
            }
            std::vector<uint32_t> ids;
            int32_t clo[3], chi[3];
            for (int a = 0; a < 3; a++) {
                clo[a] = (int32_t)std::floor(lo[a] / cs);
                chi[a] = (int32_t)std::floor(hi[a] / cs);
            }
            for (int32_t x = clo[0]; x <= chi[0]; x++)
                for (int32_t y = clo[1]; y <= chi[1]; y++)
                    for (int32_t z = clo[2]; z <= chi[2]; z++) {
                        auto it = bucket.find({x, y, z});
                        if (it == bucket.end()) continue;
                        for (uint32_t i : it->second) {
                            const float* p = xyz + 3 * (size_t)i;
                            if (p[0] >= lo[0] && p[0] <= hi[0] && p[1] >= lo[1] && p[1] <= hi[1] && p[2] >= lo[2] &&
                                p[2] <= hi[2])
                                ids.push_back(i);
                        }
                    }
            if (ids.size() < minPoints) continue;
            std::sort(ids.begin(), ids.end());
            std::vector<float> pts(3 * ids.size());
            for (size_t j = 0; j < ids.size(); j++)
                for (int a = 0; a < 3; a++) pts[3 * j + a] = xyz[3 * (size_t)ids[j] + a];
            ch.pointsBeforeReduction = ids.size();
            // 3.-4. the device part, shrinking and retrying on out-of-memory
            for (;;) {
                std::vector<float> cur = pts;
                float target = voxel / 4;
                for (int pass = 0; cur.size() / 3 > m_maxPoints; pass++) {
                    if (pass == 64) throw std::runtime_error("HipLargeScaleReconstruction: the reduction stalled");
                    cur = reduce(cur, target, 10, m_opt.reductionSeed);
                    target *= 2;
                }
                try {
                    device_chunk(ch, cur, voxel);
                    break;
                } catch (const std::runtime_error& e) {
                    if (!is_oom(e)) throw;
                    m_maxPoints = (size_t)(m_maxPoints * 0.8);
                    ch.retries++;
                    if (m_maxPoints < minPoints)
                        throw std::runtime_error("HipLargeScaleReconstruction: out of memory for chunks of " +
                                                 std::to_string(minPoints) + " points");
                }
            }
            if (ch.cells.empty()) {  // every cell erased (tcc:695-698)
                m_withoutCells.push_back(ch.coord);
                continue;
            }
            chunks.push_back(std::move(ch));
        }
        // 5. the overlap merge
        if (m_opt.mergeChunkBorders) merge(chunks, cs, overlap);
        // 6. the meshes
        for (auto& ch : chunks) mesh_chunk(ch, voxel);
        return chunks;
    }

    // maxPointsPerChunk after this object's out-of-memory shrinks
    size_t maxPointsPerChunk() const { return m_maxPoints; }

    // chunks with enough points of which the invalid erase left no cell (skipped, as the
    // reference skips a chunk whose grid has no cells, tcc:695-698)
    const std::vector<std::array<int32_t, 3>>& chunksWithoutCells() const { return m_withoutCells; }

private:
    // lattice vertex key: the integer lattice index of cell corner v (corners at index + 0 or 1)
    static uint64_t vertex_key(const int32_t* ci, int v)
    {
        const int dx = (v == 1 || v == 2 || v == 5 || v == 6) ? 1 : 0;
        const int dy = (v == 2 || v == 3 || v == 6 || v == 7) ? 1 : 0;
        const int dz = v >= 4 ? 1 : 0;
        return hip_lsr_detail::pack3(ci[0] + dx, ci[1] + dy, ci[2] + dz);
    }

    void device_chunk(HipLSRChunk& ch, const std::vector<float>& pts, float voxel)
    {
        using namespace hip_lsr_detail;
        const size_t np = pts.size() / 3;
        Handle h(pts.data(), np, m_opt.device);
        check(lvrhip_set_k(h.h, m_opt.kn, m_opt.ki, m_opt.kd));
        check(lvrhip_set_flip_point(h.h, m_opt.flipPoint[0], m_opt.flipPoint[1], m_opt.flipPoint[2]));
        check(lvrhip_calculate_normals(h.h));
        if (m_opt.ki > 0) check(lvrhip_interpolate_normals(h.h));
        std::vector<float> nrm(3 * np);
        check(lvrhip_get_normals(h.h, nrm.data()));
        size_t nv = 0, nc = 0;
        check(lvrhip_grid_build_box(h.h, voxel, m_opt.extrude ? 1 : 0, ch.gridBox.data(), &nv, &nc));
        std::vector<float> vtx(3 * nv);
        std::vector<int32_t> cells(8 * nc), ci(3 * nc);
        check(lvrhip_grid_get(h.h, vtx.data(), cells.data(), ci.data()));
        std::vector<float> proj(nv), eucl(nv);
        std::vector<uint8_t> inv(nv);
        if (nv) check(lvrhip_distances(h.h, vtx.data(), nv, voxel, proj.data(), eucl.data(), inv.data()));
        ch.points = pts;
        ch.normals = std::move(nrm);
        ch.vertices = std::move(vtx);
        ch.rawDistance = proj;
        ch.distance = std::move(proj);
        ch.invalid = std::move(inv);
        // the erase: keep the cells without an invalid corner (their ids stay the lattice's)
        ch.cells.clear();
        ch.cellIndex.clear();
        for (size_t c = 0; c < nc; c++) {
            bool bad = false;
            for (int v = 0; v < 8; v++) bad = bad || ch.invalid[cells[8 * c + v]] != 0;
            if (bad) continue;
            ch.cells.insert(ch.cells.end(), cells.begin() + 8 * c, cells.begin() + 8 * c + 8);
            ch.cellIndex.insert(ch.cellIndex.end(), ci.begin() + 3 * c, ci.begin() + 3 * c + 3);
        }
    }

    void merge(std::vector<HipLSRChunk>& chunks, float cs, float overlap)
    {
        // every chunk's (vertex key -> distance) over its remaining cells
        std::vector<std::unordered_map<uint64_t, std::pair<float, int32_t>>> qps(chunks.size());
        std::vector<std::vector<std::pair<uint64_t, int32_t>>> own(chunks.size());
        std::unordered_map<uint64_t, size_t> by_coord;
        for (size_t c = 0; c < chunks.size(); c++) {
            auto& ch = chunks[c];
            by_coord[hip_lsr_detail::pack3(ch.coord[0], ch.coord[1], ch.coord[2])] = c;
            const size_t nc = ch.cells.size() / 8;
            for (size_t k = 0; k < nc; k++)
                for (int v = 0; v < 8; v++) {
                    const uint64_t key = vertex_key(&ch.cellIndex[3 * k], v);
                    const int32_t id = ch.cells[8 * k + v];
                    if (qps[c].emplace(key, std::make_pair(ch.rawDistance[id], id)).second)
                        own[c].emplace_back(key, id);
                }
        }
        for (size_t c = 0; c < chunks.size(); c++) {
            auto& ch = chunks[c];
            for (auto& [key, id] : own[c]) {
                // the chunks holding this vertex, in chunk-coordinate order (this one included)
                std::vector<std::pair<size_t, std::pair<float, int32_t>>> holders;
                for (int dx = -1; dx <= 1; dx++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dz = -1; dz <= 1; dz++) {
                            auto it = by_coord.find(
                                hip_lsr_detail::pack3(ch.coord[0] + dx, ch.coord[1] + dy, ch.coord[2] + dz));
                            if (it == by_coord.end()) continue;
                            auto q = qps[it->second].find(key);
                            if (q != qps[it->second].end()) holders.emplace_back(it->second, q->second);
                        }
                if (holders.size() < 2) continue;
                // the position the first holder gives the vertex, so that every holder weighs alike
                const float* p = &chunks[holders[0].first].vertices[3 * (size_t)holders[0].second.second];
                double total = 0.0, wsum = 0.0;
                for (auto& [j, dv] : holders) {
                    const float d = dv.first;
                    const auto& b = chunks[j].gridBox;
                    double w = 1.0;
                    for (int a = 0; a < 3; a++) {
                        const float centre = b[a] + 0.5f * (b[3 + a] - b[a]);
                        const double fromEdge = cs / 2.0 + overlap - std::abs(centre - p[a]);
                        w = std::min(w, fromEdge / (2 * overlap));
                    }
                    w = std::max(w, 0.001);
                    total += d * w;
                    wsum += w;
                }
                ch.distance[id] = (float)(total / wsum);
            }
        }
    }

    void mesh_chunk(HipLSRChunk& ch, float voxel)
    {
        using namespace hip_lsr_detail;
        const size_t nv = ch.vertices.size() / 3;
        Handle h(ch.points.data(), ch.points.size() / 3, m_opt.device);
        size_t nv2 = 0, nc2 = 0;
        check(lvrhip_grid_build_box(h.h, voxel, m_opt.extrude ? 1 : 0, ch.gridBox.data(), &nv2, &nc2));
        if (nv2 != nv) throw std::runtime_error("HipLargeScaleReconstruction: the lattice changed between passes");
        size_t mv = 0, mf = 0;
        check(lvrhip_mesh_build_host(h.h, ch.distance.data(), ch.invalid.data(), nv, LVRHIP_MESH_MC, &mv, &mf));
        ch.meshVertices.resize(3 * mv);
        ch.meshFaces.resize(3 * mf);
        check(lvrhip_mesh_get(h.h, ch.meshVertices.data(), ch.meshFaces.data()));
    }

    HipLSROptions m_opt;
    size_t m_maxPoints;
    std::vector<std::array<int32_t, 3>> m_withoutCells;
};

}  // namespace lvr2

#endif
