// Runs lvr2::HipLargeScaleReconstruction (include/lvr2_amd/HipLargeScaleReconstruction.hpp), the
// LargeScaleReconstruction chunk loop on the device, and writes every chunk's data.
//
//   hiplsr_check run <points.f32> <n> <bgVoxelSize> <voxelSize> <kn> <ki> <kd> <fx> <fy> <fz>
//                    <mergeChunkBorders 0|1> <maxPointsPerChunk> <outdir>
//     <outdir>/chunks.txt: one line per chunk "cx cy cz points points_before_reduction retries
//     lattice_vertices cells_left mesh_vertices mesh_faces", then "empty cx cy cz" for each chunk
//     the erase left without cells, then "max_points_per_chunk N";
//     <outdir>/chunk_<i>_{box,points,normals,vertices,raw,dist}.f32, _{cells,cellindex,faces}.i32,
//     _invalid.u8, _mesh.f32 with the fields of HipLSRChunk.
#include "lvr2_amd/HipLargeScaleReconstruction.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace
{

std::vector<float> read_f32(const char* path, size_t count)
{
    std::vector<float> v(count);
    FILE* f = std::fopen(path, "rb");
    if (!f || std::fread(v.data(), sizeof(float), count, f) != count) {
        std::fprintf(stderr, "cannot read %zu floats from %s\n", count, path);
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

template <class T>
void write(const std::string& path, const std::vector<T>& v)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || (!v.empty() && std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size())) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

int run(int argc, char** argv)
{
    if (argc != 15) return 2;
    const size_t n = std::strtoull(argv[3], nullptr, 10);
    lvr2::HipLSROptions o;
    o.bgVoxelSize = std::strtof(argv[4], nullptr);
    o.voxelSize = std::strtof(argv[5], nullptr);
    o.kn = std::atoi(argv[6]);
    o.ki = std::atoi(argv[7]);
    o.kd = std::atoi(argv[8]);
    o.flipPoint = {std::strtof(argv[9], nullptr), std::strtof(argv[10], nullptr), std::strtof(argv[11], nullptr)};
    o.mergeChunkBorders = std::atoi(argv[12]) != 0;
    o.maxPointsPerChunk = std::strtoull(argv[13], nullptr, 10);
    const std::string out = argv[14];
    const auto pts = read_f32(argv[2], 3 * n);
    lvr2::HipLargeScaleReconstruction lsr(o);
    const auto chunks = lsr.reconstruct(pts.data(), n);
    FILE* f = std::fopen((out + "/chunks.txt").c_str(), "w");
    if (!f) return 2;
    for (size_t i = 0; i < chunks.size(); i++) {
        const auto& c = chunks[i];
        std::fprintf(f, "%d %d %d %zu %zu %d %zu %zu %zu %zu\n", c.coord[0], c.coord[1], c.coord[2],
                     c.points.size() / 3, c.pointsBeforeReduction, c.retries, c.vertices.size() / 3,
                     c.cells.size() / 8, c.meshVertices.size() / 3, c.meshFaces.size() / 3);
        const std::string p = out + "/chunk_" + std::to_string(i) + "_";
        write(p + "box.f32", std::vector<float>(c.gridBox.begin(), c.gridBox.end()));
        write(p + "points.f32", c.points);
        write(p + "normals.f32", c.normals);
        write(p + "vertices.f32", c.vertices);
        write(p + "cells.i32", c.cells);
        write(p + "cellindex.i32", c.cellIndex);
        write(p + "raw.f32", c.rawDistance);
        write(p + "dist.f32", c.distance);
        write(p + "invalid.u8", c.invalid);
        write(p + "mesh.f32", c.meshVertices);
        write(p + "faces.i32", c.meshFaces);
    }
    for (const auto& c : lsr.chunksWithoutCells()) std::fprintf(f, "empty %d %d %d\n", c[0], c[1], c[2]);
    std::fprintf(f, "max_points_per_chunk %zu\n", lsr.maxPointsPerChunk());
    std::fclose(f);
    std::printf("hiplsr_check run ok: %zu chunks\n", chunks.size());
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
