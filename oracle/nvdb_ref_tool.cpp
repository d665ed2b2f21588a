/*
 * Checker-side NanoVDB tool built from the reference's own sources (src/nanovdb, PNanoVDB.glsl) by
 * oracle/Makefile into oracle/_ref/nvdb_ref. TEST INFRASTRUCTURE ONLY: it writes the golden vectors that pin the
 * build's NanoVDB writer and device accessor.
 *
 *   nvdb_ref build  <dense.f32> nx ny nz dx ox oy oz <out.grid>  GridBuilder as CloudData.cpp:1383-1404
 *   nvdb_ref build16 <dense.f32> nx ny nz dx ox oy oz <out.grid> the same with GridBuilder<float, Fp16>
 *   nvdb_ref ijk    <in.grid> <ijk.i32> <out.f32>                PNanoVDB lookups at integer coordinates
 *   nvdb_ref ijk16  <in.grid> <ijk.i32> <out.f32>                the same on an Fp16 grid
 *   nvdb_ref world  <in.grid> <xyz.f32> <out.f32> <out_ijk.i32>  nearest world-space lookups
 *   nvdb_ref nvdbfile <in.grid> <out.nvdb>                      io::writeGrid of a grid (the .nvdb file format that
 *                                                                CloudData::loadFromNvdbFile reads, CloudData.cpp:1360)
 *   nvdb_ref fogsphere <radius> <voxel> <out.grid>               createFogVolumeSphere<float>(radius, 0, voxel), the
 *                                                                default scene of Module.cpp:793-795
 */
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "nanovdb/util/GridBuilder.h"
#include "nanovdb/util/IO.h"
#include "nanovdb/util/Primitives.h"

extern "C" void ref_lookup_ijk(const uint32_t* words, int n, const int32_t* ijk, float* out);
extern "C" void ref_lookup_ijk16(const uint32_t* words, int n, const int32_t* ijk, float* out);
extern "C" void ref_lookup_world(const uint32_t* words, int n, const float* xyz, float* out, int32_t* ijk_out);

template <class T>
static std::vector<T> readAll(const char* path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", path);
        std::exit(2);
    }
    std::vector<char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<T> v(b.size() / sizeof(T));
    std::memcpy(v.data(), b.data(), v.size() * sizeof(T));
    return v;
}

template <class T>
static void writeAll(const char* path, const T* data, size_t n) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(data), std::streamsize(n * sizeof(T)));
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string cmd = argv[1];
    if ((cmd == "build" || cmd == "build16") && argc == 11) {
        std::vector<float> dense = readAll<float>(argv[2]);
        const int nx = std::atoi(argv[3]), ny = std::atoi(argv[4]), nz = std::atoi(argv[5]);
        const double dx = std::atof(argv[6]);
        const nanovdb::Vec3d origin(std::atof(argv[7]), std::atof(argv[8]), std::atof(argv[9]));
        auto op = [&](const nanovdb::Coord& ijk) -> float {
            return dense[size_t(ijk.x()) + (size_t(ijk.y()) + size_t(ijk.z()) * ny) * nx];
        };
        const nanovdb::CoordBBox box(nanovdb::Coord(0), nanovdb::Coord(nx - 1, ny - 1, nz - 1));
        if (cmd == "build") {
            nanovdb::GridBuilder<float> builder(0.0f);
            builder(op, box);
            auto handle = builder.getHandle<>(dx, origin, "tmp", nanovdb::GridClass::FogVolume);
            writeAll(argv[10], handle.data(), handle.size());
        } else {
            nanovdb::GridBuilder<float, nanovdb::Fp16> builder(0.0f);
            builder(op, box);
            auto handle = builder.getHandle<>(dx, origin, "tmp", nanovdb::GridClass::FogVolume);
            writeAll(argv[10], handle.data(), handle.size());
        }
        return 0;
    }
    if (cmd == "ijk16" && argc == 5) {
        std::vector<uint32_t> grid = readAll<uint32_t>(argv[2]);
        std::vector<int32_t> ijk = readAll<int32_t>(argv[3]);
        std::vector<float> out(ijk.size() / 3);
        ref_lookup_ijk16(grid.data(), int(out.size()), ijk.data(), out.data());
        writeAll(argv[4], out.data(), out.size());
        return 0;
    }
    if (cmd == "ijk" && argc == 5) {
        std::vector<uint32_t> grid = readAll<uint32_t>(argv[2]);
        std::vector<int32_t> ijk = readAll<int32_t>(argv[3]);
        std::vector<float> out(ijk.size() / 3);
        ref_lookup_ijk(grid.data(), int(out.size()), ijk.data(), out.data());
        writeAll(argv[4], out.data(), out.size());
        return 0;
    }
    if (cmd == "world" && argc == 6) {
        std::vector<uint32_t> grid = readAll<uint32_t>(argv[2]);
        std::vector<float> xyz = readAll<float>(argv[3]);
        std::vector<float> out(xyz.size() / 3);
        std::vector<int32_t> ijk(xyz.size());
        ref_lookup_world(grid.data(), int(out.size()), xyz.data(), out.data(), ijk.data());
        writeAll(argv[4], out.data(), out.size());
        writeAll(argv[5], ijk.data(), ijk.size());
        return 0;
    }
    if (cmd == "nvdbfile" && argc == 4) {
        std::vector<uint8_t> grid = readAll<uint8_t>(argv[2]);
        auto buffer = nanovdb::HostBuffer::create(grid.size());
        std::memcpy(buffer.data(), grid.data(), grid.size());
        nanovdb::GridHandle<nanovdb::HostBuffer> handle(std::move(buffer));
        nanovdb::io::writeGrid(argv[3], handle);
        return 0;
    }
    if (cmd == "fogsphere" && argc == 5) {
        const float radius = float(std::atof(argv[2])), voxel = float(std::atof(argv[3]));
        auto handle = nanovdb::createFogVolumeSphere<float>(radius, nanovdb::Vec3<float>(0), voxel);
        writeAll(argv[4], handle.data(), handle.size());
        return 0;
    }
    std::fprintf(stderr, "usage: nvdb_ref build|build16|ijk|ijk16|world|nvdbfile|fogsphere ...\n");
    return 2;
}
