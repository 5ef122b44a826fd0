// lvr2::estimateProjectNormalsHip: estimateProjectNormals (ScanProjectUtils.cpp:361-410) with the
// batched HIP surface.  The reference loads every scan of every lidar of every scan position,
// builds an AdaptiveKSearchSurface(points, "flann", kn, ki) per scan (kd 10, calcMethod 0:
// calcPlane), sets the flip point to (0, 0, 0) of the scan's own frame, runs
// calculateSurfaceNormals (estimate + interpolation), then save() and release().
//
// Here the loaded scans go to the device in batches of at most maxBatchPoints points.  All scans
// of a batch share one handle (lvrhip_create_scans), and each is still searched on its own, as a
// surface of its own would be.  Scans with fewer than 32 * kn points, which the batched handle
// does not take, get a handle of their own.  Each scan's normals become its point buffer's normal
// array; then the scan is saved and released, in the reference's order.  Scans that are null,
// fail to load or hold no points are skipped, as the reference skips them.
//
// The project types are template parameters with lvr2's member names: p->positions[i]->lidars[j]
// ->scans[k] (shared pointers, null allowed), scan->load() / save() / release(), scan->points a
// PointBufferPtr with numPoints(), getPointArray() (get() on the result) and
// setNormalArray(floatArr, n).
#ifndef LVR2_AMD_HIPSCANPROJECT_HPP
#define LVR2_AMD_HIPSCANPROJECT_HPP

#include "../lvrhip.h"

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace lvr2
{

namespace hip_project_detail
{

inline void check(int rc)
{
    if (rc != LVRHIP_OK) throw std::runtime_error(std::string("lvrhip: ") + lvrhip_last_error());
}

struct HandleGuard {
    lvrhip_t h = nullptr;
    ~HandleGuard()
    {
        if (h) lvrhip_destroy(h);
    }
};

// kn, ki, kd 10, calcPlane, flip (0, 0, 0) of every scan; then estimate and interpolate
inline void run(lvrhip_t h, int nscans, int kn, int ki)
{
    check(lvrhip_set_k(h, kn, ki, 10));
    check(lvrhip_set_method(h, LVRHIP_METHOD_LS_PLANE));
    std::vector<float> flips(3 * (size_t)nscans, 0.0f);
    if (nscans > 1)
        check(lvrhip_set_scan_flip_points(h, flips.data()));
    else
        check(lvrhip_set_flip_point(h, 0.0f, 0.0f, 0.0f));
    check(lvrhip_calculate_normals(h));
    if (ki != 0) check(lvrhip_interpolate_normals(h));
}

}  // namespace hip_project_detail

// Returns the number of scans that got normals.
template <class ScanProjectPtrT>
size_t estimateProjectNormalsHip(ScanProjectPtrT p, size_t kn, size_t ki, size_t maxBatchPoints = (size_t)1 << 28,
                                 int device = 0)
{
    using namespace hip_project_detail;
    using ScanPtrT = typename std::decay<decltype(p->positions[0]->lidars[0]->scans[0])>::type;
    using FloatArr = typename std::decay<decltype(p->positions[0]->lidars[0]->scans[0]->points->getPointArray())>::type;
    struct Loaded {
        ScanPtrT scan;
        size_t n;
    };
    std::vector<Loaded> batch;
    size_t batch_points = 0, done = 0;
    const size_t min_batched = 32 * kn;

    auto flush = [&]() {
        if (batch.empty()) return;
        std::vector<std::vector<float>> normals(batch.size());
        std::vector<size_t> big;
        for (size_t i = 0; i < batch.size(); i++) {
            normals[i].resize(3 * batch[i].n);
            if (batch[i].n >= min_batched)
                big.push_back(i);
            else {
                // a scan too small for the batched handle: a handle of its own
                HandleGuard g;
                check(lvrhip_create(batch[i].scan->points->getPointArray().get(), batch[i].n, device, &g.h));
                run(g.h, 1, (int)kn, (int)ki);
                check(lvrhip_get_normals(g.h, normals[i].data()));
            }
        }
        if (!big.empty()) {
            std::vector<int64_t> off{0};
            std::vector<float> xyz;
            for (size_t i : big) {
                const float* src = batch[i].scan->points->getPointArray().get();
                xyz.insert(xyz.end(), src, src + 3 * batch[i].n);
                off.push_back(off.back() + (int64_t)batch[i].n);
            }
            HandleGuard g;
            if (big.size() == 1)
                check(lvrhip_create(xyz.data(), (size_t)off.back(), device, &g.h));
            else
                check(lvrhip_create_scans(xyz.data(), off.data(), (int)big.size(), device, &g.h));
            run(g.h, (int)big.size(), (int)kn, (int)ki);
            std::vector<float> all(3 * (size_t)off.back());
            check(lvrhip_get_normals(g.h, all.data()));
            for (size_t j = 0; j < big.size(); j++)
                std::memcpy(normals[big[j]].data(), all.data() + 3 * off[j], 3 * batch[big[j]].n * sizeof(float));
        }
        for (size_t i = 0; i < batch.size(); i++) {
            FloatArr arr(new float[3 * batch[i].n]);
            std::memcpy(arr.get(), normals[i].data(), 3 * batch[i].n * sizeof(float));
            batch[i].scan->points->setNormalArray(arr, batch[i].n);
            batch[i].scan->save();
            batch[i].scan->release();
            done++;
        }
        batch.clear();
        batch_points = 0;
    };

    for (size_t pos = 0; pos < p->positions.size(); pos++) {
        auto position = p->positions[pos];
        if (!position) continue;
        for (size_t l = 0; l < position->lidars.size(); l++) {
            auto lidar = position->lidars[l];
            if (!lidar) continue;
            for (size_t s = 0; s < lidar->scans.size(); s++) {
                auto scan = lidar->scans[s];
                if (!scan) continue;
                scan->load();
                if (!scan->points || scan->points->numPoints() == 0) {
                    scan->release();
                    continue;
                }
                const size_t n = scan->points->numPoints();
                if (!batch.empty() && batch_points + n > maxBatchPoints) flush();
                batch.push_back({scan, n});
                batch_points += n;
            }
        }
    }
    flush();
    return done;
}

}  // namespace lvr2

#endif
