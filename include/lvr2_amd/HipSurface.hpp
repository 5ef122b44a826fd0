// lvr2::HipSurface: header-only C++ drop-in for lvr2::CudaSurface
// (include/lvr2/reconstruction/cuda/CudaSurface.hpp:70-158) over the C ABI of liblvrhip.so
// (include/lvrhip.h).  Same method names, argument meaning and defaults (CudaSurface::init,
// CudaSurface.cu:1007-1023: kn 10, ki 10, kd 5, flip point (1e6, 1e6, 1e6)); errors are
// thrown as std::runtime_error whose message contains "out of memory" for device allocation
// failures, the substring LargeScaleReconstruction matches (LargeScaleReconstruction.tcc:656-662).
//
// calculateNormals() returns what CudaSurface::calculateNormals returns to lvr2's callers
// (lvr2_reconstruct/Main.cpp:369-370, LargeScaleReconstruction.tcc:654-673): the estimated
// normals interpolated with ki inside the same call, as CudaSurface::GPU_NN does
// (CudaSurface.cu:1135-1159); interpolateNormals() is empty there (CudaSurface.cu:1255-1257)
// and here.  The normals are those of AdaptiveKSearchSurface::calculateSurfaceNormals
// (AdaptiveKSearchSurface.tcc:158-290), not CudaSurface's approximate kd-tree window.
//
// The point and query types are template parameters so that the header compiles without the
// lvr2 headers: a points array is anything with get() (boost::shared_array<float>) or with
// width / dim / elements (LBPointArray<float>), a query point anything with
// m_position.{x,y,z}, m_distance and m_invalid (QueryPoint<BaseVector<float>>).
#ifndef LVR2_AMD_HIPSURFACE_HPP
#define LVR2_AMD_HIPSURFACE_HPP

#include "../lvrhip.h"

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace lvr2
{

class HipSurface
{
public:
    // CudaSurface(floatArr& points, size_t num_points, int device = 0)
    template <class FloatArr, class = decltype(std::declval<FloatArr&>().get())>
    HipSurface(FloatArr& points, size_t num_points, int device = 0)
    {
        create(points.get(), num_points, device);
    }

    // CudaSurface(LBPointArray<float>& points, int device = 0)
    template <class PointArray, class = decltype(std::declval<PointArray&>().elements)>
    explicit HipSurface(PointArray& points, int device = 0)
    {
        if (points.dim != 3) throw std::runtime_error("HipSurface: points must have dim 3");
        create(points.elements, points.width, device);
    }

    // Raw AoS xyz buffer (n * 3 floats, host memory).
    HipSurface(const float* xyz, size_t num_points, int device = 0) { create(xyz, num_points, device); }

    // Batched form: scans [offsets[s], offsets[s+1]) of one AoS buffer, each searched on its
    // own (estimateProjectNormals, ScanProjectUtils.cpp:361-410).
    HipSurface(const float* xyz, const std::vector<int64_t>& scan_offsets, int device = 0)
    {
        if (scan_offsets.size() < 2) throw std::runtime_error("HipSurface: scan table needs at least one scan");
        m_n = (size_t)scan_offsets.back();
        check(lvrhip_create_scans(xyz, scan_offsets.data(), (int)scan_offsets.size() - 1, device, &m_h));
    }

    HipSurface(const HipSurface&) = delete;
    HipSurface& operator=(const HipSurface&) = delete;

    ~HipSurface() { freeGPU(); }

    // CudaSurface::calculateNormals (CudaSurface.cu:1238-1253): adaptive-k normals with the
    // flip, then the interpolation with ki when ki > 0 (GPU_NN, CudaSurface.cu:1143-1148).
    void calculateNormals()
    {
        check(lvrhip_calculate_normals(m_h));
        if (m_ki > 0) check(lvrhip_interpolate_normals(m_h));
    }

    // CudaSurface::interpolateNormals is empty (CudaSurface.cu:1255-1257): calculateNormals
    // has interpolated already.
    void interpolateNormals() {}

    // getNormals(floatArr output_normals) (CudaSurface.cu:1097-1104): n * 3 floats in input order
    template <class FloatArr, class = decltype(std::declval<FloatArr&>().get())>
    void getNormals(FloatArr output_normals)
    {
        check(lvrhip_get_normals(m_h, output_normals.get()));
    }

    // getNormals(LBPointArray<float>&) (CudaSurface.cu:1083-1095): sets width and dim and
    // fills the caller's elements (n * 3 floats, allocated by the caller)
    template <class PointArray, class = decltype(std::declval<PointArray&>().elements)>
    void getNormals(PointArray& output_normals)
    {
        output_normals.dim = 3;
        output_normals.width = (decltype(output_normals.width))m_n;
        check(lvrhip_get_normals(m_h, output_normals.elements));
    }

    void getNormals(float* output_normals) { check(lvrhip_get_normals(m_h, output_normals)); }

    // Given normals (n * 3 floats, input order) instead of calculateNormals, for callers whose
    // point buffer already has normals and that only need distances
    // (LargeScaleReconstruction.tcc:640, 680-687).  No CudaSurface counterpart.
    void setNormals(const float* normals) { check(lvrhip_set_normals(m_h, normals)); }

    void setKn(int kn)
    {
        m_kn = kn;
        pushK();
    }
    void setKi(int ki)
    {
        m_ki = ki;
        pushK();
    }
    void setKd(int kd)
    {
        m_kd = kd;
        pushK();
    }

    void setFlippoint(float v_x, float v_y, float v_z) { check(lvrhip_set_flip_point(m_h, v_x, v_y, v_z)); }

    // "PCA" (calcPlaneIPCAExact), "ITERATIVE" (calcPlaneIterative), "LS" / "PLANE" (calcPlane).
    // RANSAC is not provided.
    void setMethod(const std::string& method)
    {
        int m = -1;
        if (method == "PCA") m = LVRHIP_METHOD_PCA;
        if (method == "ITERATIVE") m = LVRHIP_METHOD_ITERATIVE;
        if (method == "LS" || method == "PLANE") m = LVRHIP_METHOD_LS_PLANE;
        if (m < 0) throw std::runtime_error("HipSurface: unsupported normal method " + method);
        check(lvrhip_set_method(m_h, m));
    }

    // The points stay on the device for the handle's lifetime in either mode.
    void setReconstructionMode(bool mode = true) { m_reconstruction_mode = mode; }

    // CudaSurface::distances (CudaSurface.cu:1273-1317) with the semantics of
    // AdaptiveKSearchSurface::distance (tcc:414-457): m_distance = the signed projection
    // distance, m_invalid = eucl > 1.75 * voxel_size (PointsetGrid.tcc:138-143).
    template <class QueryPointT>
    void distances(std::vector<QueryPointT>& query_points, float voxel_size)
    {
        const size_t m = query_points.size();
        if (m == 0) return;
        std::vector<float> q(3 * m), proj(m);
        std::vector<uint8_t> inv(m);
        for (size_t i = 0; i < m; i++) {
            q[3 * i] = query_points[i].m_position.x;
            q[3 * i + 1] = query_points[i].m_position.y;
            q[3 * i + 2] = query_points[i].m_position.z;
        }
        check(lvrhip_distances(m_h, q.data(), m, voxel_size, proj.data(), nullptr, inv.data()));
        for (size_t i = 0; i < m; i++) {
            query_points[i].m_distance = proj[i];
            query_points[i].m_invalid = inv[i] != 0;
        }
    }

    // SearchTree::kSearch for many queries (idx / d2: m * k, ascending, padded with -1 / inf).
    void kSearch(const float* queries, size_t m, int k, int32_t* idx, float* d2)
    {
        check(lvrhip_knn(m_h, queries, m, k, idx, d2));
    }

    // SearchTree::radiusSearch for many queries: the k nearest with d^2 < r^2 (count: m).
    void radiusSearch(const float* queries, size_t m, int k, float r, int32_t* idx, float* d2, int32_t* count)
    {
        check(lvrhip_radius_search(m_h, queries, m, k, r, idx, d2, count));
    }

    // PointsetSurface::getInterpolatedNormal at m positions (out: m * 3 floats).
    void getInterpolatedNormals(const float* positions, size_t m, float* out)
    {
        check(lvrhip_interpolated_normals(m_h, positions, m, out));
    }

    void freeGPU()
    {
        if (m_h) {
            lvrhip_destroy(m_h);
            m_h = nullptr;
        }
    }

    size_t size() const { return m_n; }
    lvrhip_t handle() const { return m_h; }

private:
    void create(const float* xyz, size_t n, int device)
    {
        m_n = n;
        check(lvrhip_create(xyz, n, device, &m_h));
    }
    void pushK() { check(lvrhip_set_k(m_h, m_kn, m_ki, m_kd)); }
    static void check(int rc)
    {
        if (rc != LVRHIP_OK) throw std::runtime_error(std::string("lvrhip: ") + lvrhip_last_error());
    }

    lvrhip_t m_h = nullptr;
    size_t m_n = 0;
    int m_kn = 10, m_ki = 10, m_kd = 5;
    bool m_reconstruction_mode = false;
};

}  // namespace lvr2

#endif
