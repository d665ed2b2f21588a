// Host-side NanoVDB 32.3 support: header parsing of FP32 grids, extraction of the first grid of a .nvdb file,
// and the dense -> sparse conversion the reference performs with nanovdb::GridBuilder in
// CloudData::getSparseDensityField (CloudData.cpp:1367-1438).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace crvpt {

struct NanoInfo {
    uint64_t gridSize = 0;
    uint64_t rootOffset = 0;  // absolute byte offset of the root node
    uint32_t tileCount = 0;
    float background = 0.0f;
    float invMatF[9] = {0};
    float vecF[3] = {0};
    double voxelSize[3] = {1, 1, 1};
    double worldBBoxMin[3] = {0, 0, 0};
    double worldBBoxMax[3] = {0, 0, 0};
    int32_t indexBBoxMin[3] = {0, 0, 0};
    int32_t indexBBoxMax[3] = {0, 0, 0};
    uint32_t nodeCount[3] = {0, 0, 0};  // leaf, lower, upper
    uint64_t activeVoxelCount = 0;
    int valueType = 0;  // 0: FP32 leaves (GridType::Float), 1: 16-bit quantised leaves (GridType::Fp16)
};

// Validates a single-grid FP32 or Fp16 NanoVDB buffer and reads the fields the renderer needs.
bool parseNanoVdb(const uint8_t* buf, size_t size, NanoInfo& info, std::string& err);

// Parses a .xyz dense grid file (CloudData::loadFromXyzFile, CloudData.cpp:236-316): u32 sizes x, y, z, f64 voxel
// sizes, then the floats z fastest. Returns the field x fastest, min-max normalised when max - min > 1e-6 (the
// minimum seeded with 0). Sizes are validated against overflow and the file length; trailing bytes are ignored as
// the reference's stream reader does.
bool parseXyzFile(const std::vector<uint8_t>& bytes, int dims[3], double voxelSize[3], std::vector<float>& values,
                  std::string& err);

// Returns the first grid of an uncompressed .nvdb file (io::writeGrid layout: header, metadata, grid bytes).
bool extractNanoVdbGridFromFile(const std::vector<uint8_t>& file, std::vector<uint8_t>& grid, std::string& err);

// Builds a FogVolume NanoVDB grid from a dense field (x fastest) with background 0: voxels equal to the
// background stay inactive, constant fully-active leaves become lower-node tiles, voxel size dx and origin p0.
// valueType 0: FP32 leaves (nanovdb::GridBuilder<float>); 1: Fp16 leaves (nanovdb::GridBuilder<float, Fp16>,
// dithering off).
std::vector<uint8_t> buildNanoVdbFromDense(const float* values, const int dims[3], double dx, const double p0[3],
                                           const std::string& name, int valueType = 0);

// Options of the general writer behind buildNanoVdbFromDense.
struct DenseBuildOptions {
    const uint8_t* active = nullptr;  // per-voxel active flags (nullptr: a voxel is active where its value != 0)
    int32_t ijk0[3] = {0, 0, 0};      // index coordinate of values[0]; multiples of 8
    float background = 0.0f;          // root background value
    uint32_t gridClass = 0;           // nanovdb::GridClass (0 Unknown, 2 FogVolume)
};
std::vector<uint8_t> buildNanoVdb(const float* values, const int dims[3], double dx, const double p0[3],
                                  const std::string& name, int valueType, const DenseBuildOptions& opt);

// nanovdb::createFogVolumeSphere<float>(radius, center, voxelSize, halfWidth) (util/Primitives.h:84-95, 893-914):
// the level-set sphere of initSphere (distances in voxel units from the centre, a narrow band of halfWidth voxels,
// leapfrogging outside it) converted to a fog volume by GridBuilder::sdfToFog (GridBuilder.h:613-680): band voxels
// at distance v <= 0 hold v * voxelSize / (-halfWidth * voxelSize), voxels beyond the band inside the sphere 1
// (active), everything outside 0 (inactive). The root background stays the level set's, halfWidth * voxelSize.
// The reference module's default scene is createFogVolumeSphere<float>(0.25, 0, 0.01) (Module.cpp:793-795).
std::vector<uint8_t> buildFogVolumeSphere(float radius, const float center[3], double voxelSize, double halfWidth,
                                          const std::string& name);

}  // namespace crvpt
