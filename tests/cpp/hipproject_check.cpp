// Runs lvr2::estimateProjectNormalsHip (include/lvr2_amd/HipScanProject.hpp) on a scan project
// whose scans are PLY files read with include/lvr2_amd/PlyIO.hpp, with stand-ins of lvr2's scan
// project types (member names as in lvr2), and writes each scan's points and normals back as PLY.
//
//   hipproject_check run <kn> <ki> <maxBatchPoints> <outdir> <scan.ply | -> ...
//     Scan position i holds one lidar with scan i; "-" is a null scan.  A position with a null
//     lidar comes first, as the reference's loops must skip it.  Each scan that gets normals is
//     saved as <outdir>/<i>.ply (float x, y, z, nx, ny, nz).  Prints "scans with normals: N".
#include "lvr2_amd/HipScanProject.hpp"
#include "lvr2_amd/PlyIO.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace
{

// boost::shared_array<float>
struct floatArr {
    std::shared_ptr<float> p;
    floatArr() = default;
    explicit floatArr(float* a) : p(a, std::default_delete<float[]>()) {}
    float* get() const { return p.get(); }
};

struct PointBuffer {
    floatArr points, normals;
    size_t n = 0;
    size_t numPoints() const { return n; }
    floatArr getPointArray() const { return points; }
    void setNormalArray(floatArr a, size_t count)
    {
        normals = a;
        (void)count;
    }
};

// Scan (types/ScanTypes.hpp): load / save / release of the point buffer
struct Scan {
    std::string in, out;
    std::shared_ptr<PointBuffer> points;
    void load()
    {
        auto c = lvr2::hip_ply::readPly(in);
        points = std::make_shared<PointBuffer>();
        points->n = c.size();
        points->points = floatArr(new float[c.points.size() + 1]);
        std::memcpy(points->points.get(), c.points.data(), c.points.size() * sizeof(float));
    }
    void save()
    {
        lvr2::hip_ply::writePly(out, points->points.get(), points->n, points->normals.get(), nullptr, 0);
    }
    void release() { points.reset(); }
};
struct LIDAR {
    std::vector<std::shared_ptr<Scan>> scans;
};
struct ScanPosition {
    std::vector<std::shared_ptr<LIDAR>> lidars;
};
struct ScanProject {
    std::vector<std::shared_ptr<ScanPosition>> positions;
};

int run(int argc, char** argv)
{
    if (argc < 7) return 2;
    const size_t kn = std::strtoull(argv[2], nullptr, 10), ki = std::strtoull(argv[3], nullptr, 10);
    const size_t maxBatch = std::strtoull(argv[4], nullptr, 10);
    const std::string out = argv[5];
    auto project = std::make_shared<ScanProject>();
    project->positions.push_back(nullptr);
    auto with_null_lidar = std::make_shared<ScanPosition>();
    with_null_lidar->lidars.push_back(nullptr);
    project->positions.push_back(with_null_lidar);
    for (int i = 6; i < argc; i++) {
        auto pos = std::make_shared<ScanPosition>();
        auto lidar = std::make_shared<LIDAR>();
        if (std::strcmp(argv[i], "-") != 0) {
            auto scan = std::make_shared<Scan>();
            scan->in = argv[i];
            scan->out = out + "/" + std::to_string(i - 6) + ".ply";
            lidar->scans.push_back(scan);
        } else {
            lidar->scans.push_back(nullptr);
        }
        pos->lidars.push_back(lidar);
        project->positions.push_back(pos);
    }
    const size_t done = lvr2::estimateProjectNormalsHip(project, kn, ki, maxBatch);
    std::printf("scans with normals: %zu\n", done);
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
