// Runs lvr2::HipKSearchSurface (include/lvr2_amd/HipKSearchSurface.hpp) the way lvr2's
// reconstruction drives a PointsetSurface, with stand-ins of lvr2's types that carry lvr2's member
// names, and writes what it returns.
//
//   hipksearch_check run <points.f32> <n> <kn> <ki> <kd> <voxel> <lattice.f32> <nv> <cells.i32>
//                        <cell_index.i32> <nc> <queries.f32> <m> <k> <r> <outdir>
//     1. AdaptiveKSearchSurface's constructor (buffer, "FLANN", kn, ki, kd, calcMethod 3), as
//        lvr2_reconstruct's --pcm FLANN builds it (Main.cpp:291-309): the flip point it chose
//        (flip.f32), then calculateSurfaceNormals() and the point buffer's normals (normals.f32).
//     2. distance(v) for the m queries from 4 threads at once (dist_single.f32: proj, eucl per
//        query), then prepareDistances + distance(v) again, which must give the same bits.
//     3. calcDistanceValuesHip on a HashGrid stand-in built from the lattice (vertices, cells,
//        cell index, 27-neighbour links): m_distance / m_invalid per vertex (grid_dist.f32,
//        grid_inv.u8), the ids of the cells left (kept.i32); no cell left may link to an erased one.
//     4. searchTree(): kSearch and radiusSearch of each query (knn.i32 / knn_d2.f32, rad.i32 /
//        rad_d2.f32, rows of k, -1 / inf padded), kSearchParallel of all of them at once, which
//        must equal the kSearch rows.
//     5. CudaKSearchSurface's constructor (buffer, kn) with the flip point and ki set as
//        Main.cpp:330-346 does, then calculateSurfaceNormals (cuda_normals.f32).
#define LVR2_AMD_STANDALONE 1

#include <array>
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

// ---- stand-ins of the lvr2 types the header uses (names and members as in lvr2) ----------------
namespace lvr2
{

template <typename T>
struct BaseVector {
    using CoordType = T;
    T x = 0, y = 0, z = 0;
    BaseVector() = default;
    BaseVector(T a, T b, T c) : x(a), y(b), z(c) {}
};

// BoundingBox::expand / getCentroid (BoundingBox.tcc:45-66, 182-185)
template <typename BaseVecT>
struct BoundingBox {
    BaseVecT m_min{FLT_MAX, FLT_MAX, FLT_MAX}, m_max{-FLT_MAX, -FLT_MAX, -FLT_MAX}, m_centroid;
    void expand(const BaseVecT& v)
    {
        m_min = BaseVecT(std::min(v.x, m_min.x), std::min(v.y, m_min.y), std::min(v.z, m_min.z));
        m_max = BaseVecT(std::max(v.x, m_max.x), std::max(v.y, m_max.y), std::max(v.z, m_max.z));
        m_centroid = BaseVecT(m_min.x + 0.5f * (m_max.x - m_min.x), m_min.y + 0.5f * (m_max.y - m_min.y),
                              m_min.z + 0.5f * (m_max.z - m_min.z));
    }
    BaseVecT getCentroid() const { return m_centroid; }
};

// boost::shared_array<float>
struct floatArr {
    std::shared_ptr<float> p;
    floatArr() = default;
    explicit floatArr(float* a) : p(a, std::default_delete<float[]>()) {}
    float* get() const { return p.get(); }
    float& operator[](size_t i) const { return p.get()[i]; }
};

struct PointBuffer {
    floatArr points, normals;
    size_t n = 0;
    size_t numPoints() const { return n; }
    floatArr getPointArray() { return points; }
    floatArr getNormalArray() { return normals; }
    bool hasNormals() const { return normals.get() != nullptr; }
    void setNormalArray(floatArr a, size_t count)
    {
        normals = a;
        (void)count;
    }
};
using PointBufferPtr = std::shared_ptr<PointBuffer>;

// SearchTree (SearchTree.hpp:49-157)
template <typename BaseVecT>
class SearchTree
{
    using CoordT = typename BaseVecT::CoordType;

public:
    virtual ~SearchTree() = default;
    virtual int kSearch(const BaseVecT& qp, int k, std::vector<size_t>& indices,
                        std::vector<CoordT>& distances) const = 0;
    virtual int radiusSearch(const BaseVecT& qp, int k, float r, std::vector<size_t>& indices,
                             std::vector<CoordT>& distances) const = 0;
    virtual int kSearch(const BaseVecT& qp, int k, std::vector<size_t>& indices) const
    {
        std::vector<CoordT> d;
        return kSearch(qp, k, indices, d);
    }
    virtual void kSearchParallel(const BaseVecT* query, int n, int k, std::vector<size_t>& indices,
                                 std::vector<CoordT>& distances) const = 0;
};

// PointsetSurface (PointsetSurface.hpp:62-160, PointsetSurface.tcc:41-56)
template <typename BaseVecT>
class PointsetSurface
{
public:
    virtual ~PointsetSurface() = default;
    virtual std::pair<typename BaseVecT::CoordType, typename BaseVecT::CoordType> distance(BaseVecT v) const = 0;
    virtual void calculateSurfaceNormals() = 0;
    PointBufferPtr pointBuffer() const { return m_pointBuffer; }
    std::shared_ptr<SearchTree<BaseVecT>> searchTree() const { return m_searchTree; }
    const BoundingBox<BaseVecT>& getBoundingBox() const { return m_boundingBox; }
    void setKi(int k) { m_ki = k; }
    void setKn(int k) { m_kn = k; }
    void setKd(int k) { m_kd = k; }
    void setFlipPoint(const BaseVecT& p) { m_flipPoint = p; }

protected:
    explicit PointsetSurface(PointBufferPtr pointcloud) : m_pointBuffer(pointcloud)
    {
        float* p = m_pointBuffer->getPointArray().get();
        for (size_t i = 0; i < m_pointBuffer->numPoints(); i++)
            m_boundingBox.expand(BaseVecT(p[3 * i], p[3 * i + 1], p[3 * i + 2]));
    }
    PointBufferPtr m_pointBuffer;
    std::shared_ptr<SearchTree<BaseVecT>> m_searchTree;
    BoundingBox<BaseVecT> m_boundingBox;
    int m_kn = 0, m_ki = 0, m_kd = 0;
    BaseVecT m_flipPoint;
};

}  // namespace lvr2

#include "lvr2_amd/HipKSearchSurface.hpp"

namespace
{

using Vec = lvr2::BaseVector<float>;

// QueryPoint<BaseVector<float>> (QueryPoint.hpp:47-93)
struct QueryPoint {
    Vec m_position;
    float m_distance = 0;
    bool m_invalid = false;
    virtual ~QueryPoint() {}
};

// FastBox's vertex and neighbour accessors (FastBox.hpp:98-109)
struct Box {
    unsigned vertices[8];
    Box* neighbors[27] = {};
    unsigned getVertex(int i) const { return vertices[i]; }
    Box* getNeighbor(int i) const { return neighbors[i]; }
    void setNeighbor(int i, Box* b) { neighbors[i] = b; }
};

// HashGrid's query points and cell map (HashGrid.hpp:238-242, 304-310)
struct Grid {
    std::vector<QueryPoint> qps;
    std::unordered_map<int64_t, Box*> cells;
    std::vector<QueryPoint>& getQueryPoints() { return qps; }
    std::unordered_map<int64_t, Box*>& getCells() { return cells; }
    ~Grid()
    {
        for (auto& c : cells) delete c.second;
    }
};

template <class T>
std::vector<T> read_vec(const char* path, size_t count)
{
    std::vector<T> v(count);
    FILE* f = std::fopen(path, "rb");
    if (!f || std::fread(v.data(), sizeof(T), count, f) != count) {
        std::fprintf(stderr, "cannot read %zu items from %s\n", count, path);
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

void write(const std::string& path, const void* data, size_t bytes)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(data, 1, bytes, f) != bytes) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

int64_t cell_key(int x, int y, int z)
{
    return ((int64_t)(x + (1 << 20)) << 42) | ((int64_t)(y + (1 << 20)) << 21) | (int64_t)(z + (1 << 20));
}

lvr2::PointBufferPtr make_buffer(const std::vector<float>& xyz)
{
    auto buf = std::make_shared<lvr2::PointBuffer>();
    buf->n = xyz.size() / 3;
    buf->points = lvr2::floatArr(new float[xyz.size()]);
    std::memcpy(buf->points.get(), xyz.data(), xyz.size() * sizeof(float));
    return buf;
}

int run(int argc, char** argv)
{
    if (argc != 18) return 2;
    const size_t n = std::strtoull(argv[3], nullptr, 10);
    const int kn = std::atoi(argv[4]), ki = std::atoi(argv[5]), kd = std::atoi(argv[6]);
    const float voxel = std::strtof(argv[7], nullptr);
    const size_t nv = std::strtoull(argv[9], nullptr, 10), nc = std::strtoull(argv[12], nullptr, 10);
    const size_t m = std::strtoull(argv[14], nullptr, 10);
    const int k = std::atoi(argv[15]);
    const float r = std::strtof(argv[16], nullptr);
    const std::string out = argv[17];
    const auto pts = read_vec<float>(argv[2], 3 * n);
    const auto lat = read_vec<float>(argv[8], 3 * nv);
    const auto cells = read_vec<int32_t>(argv[10], 8 * nc);
    const auto cidx = read_vec<int32_t>(argv[11], 3 * nc);
    const auto q = read_vec<float>(argv[13], 3 * m);

    // 1. the AdaptiveKSearchSurface-shaped constructor and the normals
    auto buffer = make_buffer(pts);
    lvr2::HipKSearchSurface<Vec> surface(buffer, "FLANN", kn, ki, kd, 3);
    const Vec flip = surface.getBoundingBox().getCentroid();
    const float fl[3] = {flip.x, flip.y, flip.z};
    write(out + "/flip.f32", fl, sizeof(fl));
    surface.calculateSurfaceNormals();
    if (!buffer->hasNormals()) return 3;
    write(out + "/normals.f32", buffer->getNormalArray().get(), 3 * n * sizeof(float));

    // 2. distance(v) from several threads, then from the prepared batch
    std::vector<float> single(2 * m), cached(2 * m);
    {
        std::vector<std::thread> ts;
        for (int t = 0; t < 4; t++)
            ts.emplace_back([&, t] {
                for (size_t i = t; i < m; i += 4) {
                    auto d = surface.distance(Vec(q[3 * i], q[3 * i + 1], q[3 * i + 2]));
                    single[2 * i] = d.first;
                    single[2 * i + 1] = d.second;
                }
            });
        for (auto& t : ts) t.join();
    }
    std::vector<QueryPoint> qps(m);
    for (size_t i = 0; i < m; i++) qps[i].m_position = Vec(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
    surface.prepareDistances(qps);
    for (size_t i = 0; i < m; i++) {
        auto d = surface.distance(qps[i].m_position);
        cached[2 * i] = d.first;
        cached[2 * i + 1] = d.second;
    }
    write(out + "/dist_single.f32", single.data(), single.size() * sizeof(float));
    if (std::memcmp(single.data(), cached.data(), single.size() * sizeof(float)) != 0) return 4;

    // 3. calcDistanceValuesHip on the grid stand-in
    {
        Grid grid;
        grid.qps.resize(nv);
        for (size_t i = 0; i < nv; i++) grid.qps[i].m_position = Vec(lat[3 * i], lat[3 * i + 1], lat[3 * i + 2]);
        std::vector<Box*> boxes(nc);
        for (size_t c = 0; c < nc; c++) {
            boxes[c] = new Box;
            for (int v = 0; v < 8; v++) boxes[c]->vertices[v] = (unsigned)cells[8 * c + v];
            grid.cells[(int64_t)c] = boxes[c];
        }
        std::unordered_map<int64_t, Box*> by_index;
        for (size_t c = 0; c < nc; c++) by_index[cell_key(cidx[3 * c], cidx[3 * c + 1], cidx[3 * c + 2])] = boxes[c];
        for (size_t c = 0; c < nc; c++)
            for (int d = 0; d < 27; d++) {
                auto it = by_index.find(cell_key(cidx[3 * c] + d / 9 - 1, cidx[3 * c + 1] + (d / 3) % 3 - 1,
                                                 cidx[3 * c + 2] + d % 3 - 1));
                if (it != by_index.end() && d != 13) boxes[c]->neighbors[d] = it->second;
            }
        lvr2::calcDistanceValuesHip(grid, surface, voxel);
        std::vector<float> gd(nv);
        std::vector<uint8_t> gi(nv);
        for (size_t i = 0; i < nv; i++) {
            gd[i] = grid.qps[i].m_distance;
            gi[i] = grid.qps[i].m_invalid ? 1 : 0;
        }
        std::vector<int32_t> kept;
        for (auto& c : grid.cells) kept.push_back((int32_t)c.first);
        std::sort(kept.begin(), kept.end());
        // a cell left links only to cells left
        std::unordered_map<const Box*, bool> alive;
        for (auto& c : grid.cells) alive[c.second] = true;
        for (auto& c : grid.cells)
            for (int d = 0; d < 27; d++)
                if (c.second->neighbors[d] && !alive.count(c.second->neighbors[d])) return 5;
        write(out + "/grid_dist.f32", gd.data(), nv * sizeof(float));
        write(out + "/grid_inv.u8", gi.data(), nv);
        write(out + "/kept.i32", kept.data(), kept.size() * sizeof(int32_t));
    }

    // 4. the search tree
    {
        auto tree = surface.searchTree();
        std::vector<int32_t> ki_(m * k, -1), ri(m * k, -1);
        std::vector<float> kd2(m * k, INFINITY), rd2(m * k, INFINITY);
        std::vector<size_t> idx;
        std::vector<float> d;
        for (size_t i = 0; i < m; i++) {
            const Vec p(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
            const int nk = tree->kSearch(p, k, idx, d);
            for (int j = 0; j < nk; j++) {
                ki_[i * k + j] = (int32_t)idx[j];
                kd2[i * k + j] = d[j];
            }
            std::vector<size_t> idx2;
            tree->kSearch(p, k, idx2);
            if (idx2 != idx) return 6;
            const int nr = tree->radiusSearch(p, k, r, idx, d);
            for (int j = 0; j < nr; j++) {
                ri[i * k + j] = (int32_t)idx[j];
                rd2[i * k + j] = d[j];
            }
        }
        std::vector<Vec> qv(m);
        for (size_t i = 0; i < m; i++) qv[i] = Vec(q[3 * i], q[3 * i + 1], q[3 * i + 2]);
        std::vector<size_t> pidx;
        std::vector<float> pd2;
        tree->kSearchParallel(qv.data(), (int)m, k, pidx, pd2);
        for (size_t j = 0; j < m * (size_t)k; j++) {
            const int32_t a = pidx[j] == SIZE_MAX ? -1 : (int32_t)pidx[j];
            if (a != ki_[j] || std::memcmp(&pd2[j], &kd2[j], 4) != 0) return 7;
        }
        write(out + "/knn.i32", ki_.data(), ki_.size() * sizeof(int32_t));
        write(out + "/knn_d2.f32", kd2.data(), kd2.size() * sizeof(float));
        write(out + "/rad.i32", ri.data(), ri.size() * sizeof(int32_t));
        write(out + "/rad_d2.f32", rd2.data(), rd2.size() * sizeof(float));
    }

    // 5. the CudaKSearchSurface-shaped constructor, set up as Main.cpp:330-346 does
    {
        auto buffer2 = make_buffer(pts);
        lvr2::HipKSearchSurface<Vec> cuda_like(buffer2, (size_t)kn);
        cuda_like.setKd(kd);
        cuda_like.setKi(ki);
        cuda_like.setKn(kn);
        cuda_like.setFlipPoint(Vec(fl[0], fl[1], fl[2]));
        cuda_like.calculateSurfaceNormals();
        write(out + "/cuda_normals.f32", buffer2->getNormalArray().get(), 3 * n * sizeof(float));
    }
    std::printf("hipksearch_check run ok: %zu points, %zu lattice vertices, %zu cells, %zu queries\n", n, nv, nc, m);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    try {
        if (std::strcmp(argv[1], "run") == 0) return run(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 11;
    }
    return 2;
}
