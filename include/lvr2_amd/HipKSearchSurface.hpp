// lvr2::HipKSearchSurface<BaseVecT>: the PointsetSurface-side GPU surface, over the C ABI of
// liblvrhip.so (include/lvrhip.h).  It takes the place of
//   - CudaKSearchSurface<BaseVecT> (CudaKSearchSurface.hpp, CudaKSearchSurface.tcc:12-68), the
//     surface lvr2_reconstruct builds for --pcm LBVH_CUDA (lvr2_reconstruct/Main.cpp:310-321), and
//   - AdaptiveKSearchSurface<BaseVecT> (AdaptiveKSearchSurface.tcc:56-156) for --pcm FLANN / LVR2
//     (Main.cpp:291-309),
// behind the PointsetSurface interface (PointsetSurface.hpp:62-160) that PointsetGrid,
// FastReconstruction and BilinearFastBox call:
//   - the constructors take the reference's arguments; init() puts the flip point at the bounding
//     box centroid (AdaptiveKSearchSurface.tcc:141-156; BoundingBox.tcc:62-66 computes it in fp32 as
//     min + 0.5f * (max - min));
//   - calculateSurfaceNormals() estimates the normals on the device and interpolates them with ki
//     when ki != 0 (AdaptiveKSearchSurface.tcc:158-290), then stores them as the point buffer's
//     normal array, as the reference does (tcc:166-167);
//   - distance(v) const is the Hoppe signed distance of AdaptiveKSearchSurface::distance
//     (tcc:414-457).  It is thread-safe.  A position that prepareDistances() has seen is answered
//     from that one batched device call, so PointsetGrid::calcDistanceValues's OpenMP loop
//     (PointsetGrid.tcc:111-150) runs unchanged after one prepareDistances(grid.getQueryPoints());
//     any other position costs one device call;
//   - searchTree() is a SearchTree (SearchTree.hpp:49-157) over the same device LBVH: kSearch,
//     radiusSearch (the k nearest with d^2 < r^2, as SearchKDTree::radiusSearch,
//     KDTree.tcc:368-386) and a batched kSearchParallel.
// calcDistanceValuesHip(grid, surface, voxelsize) below is PointsetGrid::calcDistanceValues
// (PointsetGrid.tcc:111-185) with the batched call in place of the per-vertex loop.
//
// Every search tree name maps to the device LBVH, which returns the exact k nearest points under
// the fp32 key (d^2, index); "flann" (SearchTreeFlann over a KDTreeSingleIndex, an exact search)
// and "lvr2" (SearchKDTree) are the reference's exact trees.  calcMethod: 0 calcPlane (least
// squares), 2 calcPlaneIterative, 3 calcPlaneIPCAExact (the reference's switch, tcc:225-246);
// 1 (RANSAC) and a scan pose file are not provided and throw std::runtime_error.
//
// Built inside lvr2, the header includes lvr2's PointsetSurface and SearchTree.  Defining
// LVR2_AMD_STANDALONE before including it skips those includes: the includer then supplies
// lvr2::PointsetSurface, lvr2::SearchTree, lvr2::PointBufferPtr and lvr2::floatArr with lvr2's
// member names (the C++ test program does this).
#ifndef LVR2_AMD_HIPKSEARCHSURFACE_HPP
#define LVR2_AMD_HIPKSEARCHSURFACE_HPP

#ifndef LVR2_AMD_STANDALONE
#include "lvr2/reconstruction/PointsetSurface.hpp"
#include "lvr2/reconstruction/SearchTree.hpp"
#endif

#include "HipSurface.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace lvr2
{

namespace hip_detail
{

// One device handle, shared by a surface and its search tree; the mutex serialises its calls.
struct SharedHandle {
    std::unique_ptr<HipSurface> hs;
    std::mutex mu;
};

// A position's exact bit pattern, the key of the prepared distances.
struct PosKey {
    uint32_t b[3];
    bool operator==(const PosKey& o) const { return b[0] == o.b[0] && b[1] == o.b[1] && b[2] == o.b[2]; }
};
struct PosHash {
    size_t operator()(const PosKey& k) const
    {
        uint64_t h = 0x9E3779B97F4A7C15ull;
        for (uint32_t v : k.b) h = (h ^ v) * 0xBF58476D1CE4E5B9ull;
        return (size_t)(h ^ (h >> 31));
    }
};
inline PosKey pos_key(float x, float y, float z)
{
    PosKey k;
    std::memcpy(&k.b[0], &x, 4);
    std::memcpy(&k.b[1], &y, 4);
    std::memcpy(&k.b[2], &z, 4);
    return k;
}

}  // namespace hip_detail

// SearchTree over the device LBVH (SearchTree.hpp:49-157).
template <typename BaseVecT>
class HipSearchTree : public SearchTree<BaseVecT>
{
    using CoordT = typename BaseVecT::CoordType;

public:
    explicit HipSearchTree(std::shared_ptr<hip_detail::SharedHandle> dev) : m_dev(std::move(dev)) {}

    // The k nearest points, ascending; fewer when the cloud has fewer.  distances: squared.
    int kSearch(const BaseVecT& qp, int k, std::vector<size_t>& indices,
                std::vector<CoordT>& distances) const override
    {
        return rows(qp, k, -1.0f, indices, distances);
    }
    using SearchTree<BaseVecT>::kSearch;

    // The k nearest points with d^2 < r^2 (SearchKDTree::radiusSearch, KDTree.tcc:368-386).
    int radiusSearch(const BaseVecT& qp, int k, float r, std::vector<size_t>& indices,
                     std::vector<CoordT>& distances) const override
    {
        return rows(qp, k, r, indices, distances);
    }

    // n queries in one device call; row i fills indices / distances [i * k, i * k + k) (slots past
    // the cloud's size: SIZE_MAX / +inf).
    void kSearchParallel(const BaseVecT* query, int n, int k, std::vector<size_t>& indices,
                         std::vector<CoordT>& distances) const override
    {
        if (n <= 0 || k <= 0) return;
        std::vector<float> q(3 * (size_t)n), d2((size_t)n * k);
        std::vector<int32_t> idx((size_t)n * k);
        for (int i = 0; i < n; i++) {
            q[3 * (size_t)i] = query[i].x;
            q[3 * (size_t)i + 1] = query[i].y;
            q[3 * (size_t)i + 2] = query[i].z;
        }
        {
            std::lock_guard<std::mutex> g(m_dev->mu);
            m_dev->hs->kSearch(q.data(), (size_t)n, k, idx.data(), d2.data());
        }
        if (indices.size() < (size_t)n * k) indices.resize((size_t)n * k);
        if (distances.size() < (size_t)n * k) distances.resize((size_t)n * k);
        for (size_t j = 0; j < (size_t)n * k; j++) {
            indices[j] = idx[j] >= 0 ? (size_t)idx[j] : std::numeric_limits<size_t>::max();
            distances[j] = (CoordT)d2[j];
        }
    }

private:
    int rows(const BaseVecT& qp, int k, float r, std::vector<size_t>& indices, std::vector<CoordT>& distances) const
    {
        indices.clear();
        distances.clear();
        if (k <= 0) return 0;
        const float q[3] = {(float)qp.x, (float)qp.y, (float)qp.z};
        std::vector<int32_t> idx(k);
        std::vector<float> d2(k);
        {
            std::lock_guard<std::mutex> g(m_dev->mu);
            if (r < 0.0f)
                m_dev->hs->kSearch(q, 1, k, idx.data(), d2.data());
            else
                m_dev->hs->radiusSearch(q, 1, k, r, idx.data(), d2.data(), nullptr);
        }
        for (int j = 0; j < k && idx[j] >= 0; j++) {
            indices.push_back((size_t)idx[j]);
            distances.push_back((CoordT)d2[j]);
        }
        return (int)indices.size();
    }

    std::shared_ptr<hip_detail::SharedHandle> m_dev;
};

template <typename BaseVecT>
class HipKSearchSurface : public PointsetSurface<BaseVecT>
{
    using CoordT = typename BaseVecT::CoordType;

public:
    // CudaKSearchSurface(PointBufferPtr pbuffer, size_t k) (CudaKSearchSurface.tcc:12-29): kn = k;
    // ki and kd start at AdaptiveKSearchSurface's defaults (10), the caller sets them as
    // Main.cpp:330-332 does.  The plane fit is calcPlaneIPCAExact, the smallest eigenvector of the
    // neighbourhood covariance (CudaKSearchSurface's knn_normals kernel fits its plane to the same
    // covariance by a weighted-determinant method, knn_normals_kernel.cu:90-160).
    HipKSearchSurface(PointBufferPtr pbuffer, size_t k, int device = 0)
        : HipKSearchSurface(pbuffer, std::string("lbvh_cuda"), (int)k, 10, 10, 3, std::string(), device)
    {
    }

    // AdaptiveKSearchSurface(buffer, searchTreeName, kn, ki, kd, calcMethod, posefile)
    // (AdaptiveKSearchSurface.hpp:99-107, tcc:56-91)
    HipKSearchSurface(PointBufferPtr buffer, std::string searchTreeName, int kn = 10, int ki = 10, int kd = 10,
                      int calcMethod = 0, std::string posefile = std::string(), int device = 0)
        : PointsetSurface<BaseVecT>(buffer), m_calcMethod(calcMethod)
    {
        std::transform(searchTreeName.begin(), searchTreeName.end(), searchTreeName.begin(),
                       [](unsigned char c) { return (char)std::tolower(c); });
        m_searchTreeName = searchTreeName;
        if (calcMethod == 1) throw std::runtime_error("HipKSearchSurface: RANSAC normals (calcMethod 1) are not provided");
        if (!posefile.empty()) throw std::runtime_error("HipKSearchSurface: scan pose files are not provided");
        this->setKi(ki);
        this->setKn(kn);
        this->setKd(kd);
        init();
        m_dev = std::make_shared<hip_detail::SharedHandle>();
        const size_t n = this->m_pointBuffer->numPoints();
        auto pts = this->m_pointBuffer->getPointArray();
        m_dev->hs.reset(new HipSurface(pts.get(), n, device));
        m_dev->hs->setReconstructionMode(true);
        m_dev->hs->setMethod(calcMethod == 2 ? "ITERATIVE" : calcMethod == 3 ? "PCA" : "LS");
        this->m_searchTree = std::make_shared<HipSearchTree<BaseVecT>>(m_dev);
    }

    virtual ~HipKSearchSurface() {}

    // AdaptiveKSearchSurface::init (tcc:141-156): the flip point is the bounding box centroid.
    void init() { this->m_flipPoint = this->m_boundingBox.getCentroid(); }

    // AdaptiveKSearchSurface::calculateSurfaceNormals (tcc:158-290): estimate, flip towards the
    // flip point, interpolate with ki when ki != 0; the result becomes the point buffer's normals.
    void calculateSurfaceNormals() override
    {
        std::lock_guard<std::mutex> g(m_dev->mu);
        push();
        m_dev->hs->calculateNormals();
        const size_t n = this->m_pointBuffer->numPoints();
        floatArr normals(new float[3 * n]);
        m_dev->hs->getNormals(normals.get());
        this->m_pointBuffer->setNormalArray(normals, n);
        m_has_normals = true;
        m_prepared.clear();
    }

    // AdaptiveKSearchSurface::distance (tcc:414-457): (projected distance, euclidean distance).
    std::pair<CoordT, CoordT> distance(BaseVecT v) const override
    {
        const float x = (float)v.x, y = (float)v.y, z = (float)v.z;
        {
            std::lock_guard<std::mutex> g(m_dev->mu);
            if (!m_prepared.empty()) {
                auto it = m_prepared.find(hip_detail::pos_key(x, y, z));
                if (it != m_prepared.end()) return {(CoordT)it->second.first, (CoordT)it->second.second};
            }
            push_normals();
            const float q[3] = {x, y, z};
            float proj = 0.0f, eucl = 0.0f;
            check(lvrhip_distances(m_dev->hs->handle(), q, 1, 1.0f, &proj, &eucl, nullptr));
            return {(CoordT)proj, (CoordT)eucl};
        }
    }

    // One batched device call for every position of query_points (anything with m_position):
    // distance() then answers these positions without a device call.
    template <class QueryPointT>
    void prepareDistances(const std::vector<QueryPointT>& query_points)
    {
        const size_t m = query_points.size();
        std::vector<float> q(3 * m), proj(m), eucl(m);
        for (size_t i = 0; i < m; i++) {
            q[3 * i] = (float)query_points[i].m_position.x;
            q[3 * i + 1] = (float)query_points[i].m_position.y;
            q[3 * i + 2] = (float)query_points[i].m_position.z;
        }
        std::lock_guard<std::mutex> g(m_dev->mu);
        push_normals();
        if (m) check(lvrhip_distances(m_dev->hs->handle(), q.data(), m, 1.0f, proj.data(), eucl.data(), nullptr));
        m_prepared.clear();
        m_prepared.reserve(m);
        for (size_t i = 0; i < m; i++)
            m_prepared[hip_detail::pos_key(q[3 * i], q[3 * i + 1], q[3 * i + 2])] = {proj[i], eucl[i]};
    }

    // m_distance and m_invalid (eucl > 1.75 voxel, PointsetGrid.tcc:138-143) of every query point
    // in one batched device call.
    template <class QueryPointT>
    void distances(std::vector<QueryPointT>& query_points, float voxelsize)
    {
        std::lock_guard<std::mutex> g(m_dev->mu);
        push_normals();
        m_dev->hs->distances(query_points, voxelsize);
    }

    // The surface's device handle (for the C ABI calls the class does not wrap: the lattice,
    // the mesh, radius search of many queries).
    HipSurface& device() { return *m_dev->hs; }

private:
    // PointsetSurface::setKn / setKi / setKd / setFlipPoint are not virtual: the values are sent
    // to the device before each call that uses them.
    void push() const
    {
        m_dev->hs->setKn(this->m_kn);
        m_dev->hs->setKi(this->m_ki);
        m_dev->hs->setKd(this->m_kd);
        m_dev->hs->setFlippoint((float)this->m_flipPoint.x, (float)this->m_flipPoint.y, (float)this->m_flipPoint.z);
    }
    // A point buffer that came with normals (Main.cpp:350 skips the estimation then): those.
    void push_normals() const
    {
        push();
        if (m_has_normals) return;
        if (!this->m_pointBuffer->hasNormals())
            throw std::runtime_error("HipKSearchSurface: distance needs normals (calculateSurfaceNormals first)");
        auto nrm = this->m_pointBuffer->getNormalArray();
        m_dev->hs->setNormals(nrm.get());
        m_has_normals = true;
    }
    static void check(int rc)
    {
        if (rc != LVRHIP_OK) throw std::runtime_error(std::string("lvrhip: ") + lvrhip_last_error());
    }

    std::string m_searchTreeName;
    int m_calcMethod;
    std::shared_ptr<hip_detail::SharedHandle> m_dev;
    mutable bool m_has_normals = false;
    std::unordered_map<hip_detail::PosKey, std::pair<float, float>, hip_detail::PosHash> m_prepared;
};

// PointsetGrid::calcDistanceValues (PointsetGrid.tcc:111-185) with one batched device call in
// place of the per-vertex OpenMP loop: every query point's m_distance and m_invalid, then every
// cell with an invalid corner is unlinked from its 27 neighbours, deleted and erased from the
// grid's cell map.  GridT: getQueryPoints() and getCells() as HashGrid (HashGrid.hpp:238-242).
template <class GridT, class SurfaceT>
void calcDistanceValuesHip(GridT& grid, SurfaceT& surface, float voxelsize)
{
    auto& qps = grid.getQueryPoints();
    surface.distances(qps, voxelsize);
    auto& cells = grid.getCells();
    for (auto it = cells.begin(); it != cells.end();) {
        auto* box = it->second;
        bool invalid = false;
        for (int v = 0; v < 8 && !invalid; v++) invalid = qps[box->getVertex(v)].m_invalid;
        if (!invalid) {
            ++it;
            continue;
        }
        for (int k = 0; k < 27; k++) {
            auto* nb = box->getNeighbor(k);
            if (nb != nullptr) nb->setNeighbor(26 - k, nullptr);
        }
        delete box;
        it = cells.erase(it);
    }
}

}  // namespace lvr2

#endif
