// NanoVDB 32.3 writer and reader for FP32 and Fp16 fog volumes (see nanovdb_build.hpp).
//
// Serialised layout, breadth first (every node 32-byte aligned, offsets from the node that holds them):
//   GridData 672 B | TreeData 64 B | RootData 64 B + 32 B per tile | upper nodes 270400 B | lower nodes 33856 B |
//   leaves 2144 B (FP32 values) or 1120 B (Fp16: per-leaf minimum and quantum, 16-bit codes).
// Root tiles are ordered by tile origin (x, then y, then z), children of an internal node by ascending mask bit,
// which is the order nanovdb::GridBuilder emits them in.
#include "nanovdb_build.hpp"

#include "parallel_for.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

namespace crvpt {
namespace {

constexpr uint64_t kMagic = 0x304244566f6e614eULL;  // "NanoVDB0"
constexpr uint32_t kVersion = (32u << 21) | (3u << 10) | 3u;
constexpr uint32_t kGridSize = 672, kTreeSize = 64, kRootSize = 64, kTileSize = 32;
constexpr uint32_t kUpperSize = 270400, kLowerSize = 33856, kLeafSize = 2144, kLeafSizeFp16 = 1120;
constexpr uint32_t kGridTypeFloat = 1, kGridTypeFp16 = 15;

template <class T>
void put(std::vector<uint8_t>& b, size_t off, T v) {
    std::memcpy(b.data() + off, &v, sizeof(T));
}
template <class T>
T get(const uint8_t* b, size_t off) {
    T v;
    std::memcpy(&v, b + off, sizeof(T));
    return v;
}

struct Coord {
    int x, y, z;
    bool operator<(const Coord& o) const { return std::tie(x, y, z) < std::tie(o.x, o.y, o.z); }
};

struct Leaf {
    Coord origin;
    float values[512];
    uint64_t mask[8];
    uint64_t offset = 0;
};
struct Lower {
    Coord origin;
    uint64_t childMask[64] = {0}, valueMask[64] = {0};
    float tileValue[4096];
    int child[4096];  // leaf index or -1
    uint64_t offset = 0;
};
struct Upper {
    Coord origin;
    uint64_t childMask[512] = {0}, valueMask[512] = {0};
    float tileValue[32768];
    int child[32768];  // lower index or -1
    uint64_t offset = 0;
};

// Running statistics in the incremental form nanovdb::Stats uses (GridStats.h:243-297): min, max, mean and the
// aggregated squared deviation, so node averages and deviations round exactly as the reference's.
struct Stats {
    int32_t bmin[3] = {INT32_MAX, INT32_MAX, INT32_MAX};
    int32_t bmax[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    float mn = 3.402823466e38f, mx = -3.402823466e38f;
    uint64_t count = 0;
    double avgD = 0.0, aux = 0.0;
    bool empty() const { return bmin[0] > bmax[0] || bmin[1] > bmax[1] || bmin[2] > bmax[2]; }
    void extrema(float v) {
        mn = std::min(mn, v);
        mx = std::max(mx, v);
    }
    void addValue(float v) {
        extrema(v);
        count += 1;
        const double delta = double(v) - avgD;
        avgD += delta / double(count);
        aux += delta * (double(v) - avgD);
    }
    void addValue(float v, uint64_t n) {
        const double denom = 1.0 / double(count + n);
        const double delta = double(v) - avgD;
        avgD += denom * delta * double(n);
        aux += denom * delta * delta * double(count) * double(n);
        extrema(v);
        count += n;
    }
    void addStats(const Stats& o) {
        if (o.count > 0) {
            const double denom = 1.0 / double(count + o.count);
            const double delta = o.avgD - avgD;
            avgD += denom * delta * double(o.count);
            aux += o.aux + denom * delta * delta * double(count) * double(o.count);
            extrema(o.mn);
            extrema(o.mx);
            count += o.count;
        }
    }
    void addBox(const int32_t lo[3], const int32_t hi[3]) {
        for (int a = 0; a < 3; a++) {
            bmin[a] = std::min(bmin[a], lo[a]);
            bmax[a] = std::max(bmax[a], hi[a]);
        }
    }
    void merge(const Stats& o) {  // NodeStats::add: bounding box union and statistics
        if (!o.empty()) addBox(o.bmin, o.bmax);
        addStats(o);
    }
    float avg() const { return float(avgD); }
    float stddev() const { return float(std::sqrt(count < 2 ? 0.0 : aux / double(count))); }
};

// nanovdb::Map::applyMap for the scale-translate map: fma(x, dx, T) per axis (NanoVDB.h:1433-1438).
inline void applyMap(double dx, const double p0[3], const double in[3], double out[3]) {
    for (int a = 0; a < 3; a++) out[a] = std::fma(in[a], dx, std::fma(0.0, 0.0, std::fma(0.0, 0.0, p0[a])));
}

// CRC-32 (reflected polynomial 0xEDB88320, as in nanovdb::CRC32, GridChecksum.h:62-96).
uint32_t crc32(const uint8_t* data, size_t n) {
    static uint32_t lut[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++) c = (c >> 1) ^ ((c & 1u) ? 0xEDB88320u : 0u);
            lut[i] = c;
        }
        init = true;
    }
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; i++) c = lut[(c ^ data[i]) & 0xFFu] ^ (c >> 8);
    return ~c;
}

void putCoord(std::vector<uint8_t>& b, size_t off, const int32_t c[3]) {
    for (int a = 0; a < 3; a++) put<int32_t>(b, off + 4 * a, c[a]);
}

}  // namespace

bool parseNanoVdb(const uint8_t* buf, size_t size, NanoInfo& info, std::string& err) {
    if (size < kGridSize + kTreeSize + kRootSize) {
        err = "NanoVDB buffer too small";
        return false;
    }
    if (get<uint64_t>(buf, 0) != kMagic) {
        err = "NanoVDB buffer: bad magic number";
        return false;
    }
    if ((get<uint32_t>(buf, 16) >> 21) != 32u) {
        err = "NanoVDB buffer: unsupported major version (expected 32)";
        return false;
    }
    // CloudData::computeSparseGridMetadata rejects non-float grids (CloudData.cpp:1292-1296); this reader also
    // takes NanoVDB's 16-bit quantised float grids.
    const uint32_t gridType = get<uint32_t>(buf, 636);
    if (gridType != kGridTypeFloat && gridType != kGridTypeFp16) {
        err = "Fatal error in CloudData::computeSparseGridMetadata: The grid handle does not store a grid with value "
              "type float (or Fp16).";
        return false;
    }
    info.valueType = gridType == kGridTypeFp16 ? 1 : 0;
    info.gridSize = get<uint64_t>(buf, 32);
    if (info.gridSize > size) {
        err = "NanoVDB buffer: grid size exceeds the buffer";
        return false;
    }
    for (int i = 0; i < 9; i++) info.invMatF[i] = get<float>(buf, 296 + 36 + 4 * i);
    for (int i = 0; i < 3; i++) info.vecF[i] = get<float>(buf, 296 + 72 + 4 * i);
    for (int i = 0; i < 3; i++) {
        info.worldBBoxMin[i] = get<double>(buf, 560 + 8 * i);
        info.worldBBoxMax[i] = get<double>(buf, 584 + 8 * i);
        info.voxelSize[i] = get<double>(buf, 608 + 8 * i);
    }
    const uint64_t rootRel = get<uint64_t>(buf, kGridSize + 24);
    info.rootOffset = kGridSize + rootRel;
    for (int i = 0; i < 3; i++) info.nodeCount[i] = get<uint32_t>(buf, kGridSize + 32 + 4 * i);
    info.activeVoxelCount = get<uint64_t>(buf, kGridSize + 56);
    if (info.rootOffset + kRootSize > size) {
        err = "NanoVDB buffer: root outside the buffer";
        return false;
    }
    for (int i = 0; i < 3; i++) {
        info.indexBBoxMin[i] = get<int32_t>(buf, info.rootOffset + 4 * i);
        info.indexBBoxMax[i] = get<int32_t>(buf, info.rootOffset + 12 + 4 * i);
    }
    info.tileCount = get<uint32_t>(buf, info.rootOffset + 24);
    info.background = get<float>(buf, info.rootOffset + 28);
    if (info.rootOffset + kRootSize + uint64_t(kTileSize) * info.tileCount > size) {
        err = "NanoVDB buffer: root tiles outside the buffer";
        return false;
    }
    return true;
}

bool parseXyzFile(const std::vector<uint8_t>& bytes, int dims[3], double voxelSize[3], std::vector<float>& values,
                  std::string& err) {
    if (bytes.size() < 36) {
        err = "load_cloud_file: truncated .xyz header";
        return false;
    }
    uint32_t sz[3];
    std::memcpy(sz, bytes.data(), 12);
    std::memcpy(voxelSize, bytes.data() + 12, 24);
    // the header is untrusted: empty grids, size products that overflow and sizes beyond int are rejected
    size_t n = 0, payload = 0;
    if (sz[0] == 0 || sz[1] == 0 || sz[2] == 0 || sz[0] > 0x7FFFFFFFu || sz[1] > 0x7FFFFFFFu || sz[2] > 0x7FFFFFFFu ||
        __builtin_mul_overflow(size_t(sz[0]), size_t(sz[1]), &n) || __builtin_mul_overflow(n, size_t(sz[2]), &n) ||
        __builtin_mul_overflow(n, sizeof(float), &payload) || payload > bytes.size() - 36) {
        err = "load_cloud_file: truncated or invalid .xyz data";
        return false;
    }
    std::vector<float> src(n);
    std::memcpy(src.data(), bytes.data() + 36, payload);
    values.assign(n, 0.0f);
    const size_t X = sz[0], Y = sz[1], Z = sz[2];
    parallelFor(Z, [&](size_t z0, size_t z1, size_t) {
        for (size_t z = z0; z < z1; z++)
            for (size_t y = 0; y < Y; y++)
                for (size_t x = 0; x < X; x++) values[x + (y + z * Y) * X] = src[z + (y + x * Y) * Z];
    }, 1);
    float mn = 0.0f, mx = -3.402823466e38f;  // sgl::reduceFloatArrayMinMax seeded with (0, lowest)
    for (float v : values) {
        mn = std::min(mn, v);
        mx = std::max(mx, v);
    }
    if (mx - mn > 1e-6f) {
        parallelFor(n, [&](size_t b, size_t e, size_t) {
            for (size_t i = b; i < e; i++) values[i] = (values[i] - mn) / (mx - mn);
        }, 1 << 20);
    }
    for (int a = 0; a < 3; a++) dims[a] = int(sz[a]);
    return true;
}

bool extractNanoVdbGridFromFile(const std::vector<uint8_t>& file, std::vector<uint8_t>& grid, std::string& err) {
    // nanovdb::io segment: Header (16 B: magic, version, grid count u16, codec u16), then per grid a 176-byte
    // MetaData record followed by the grid name, then the grid buffers.
    if (file.size() < 16 || get<uint64_t>(file.data(), 0) != kMagic) {
        err = "load_cloud_file: not a NanoVDB file";
        return false;
    }
    const uint16_t gridCount = get<uint16_t>(file.data(), 12);
    const uint16_t codec = get<uint16_t>(file.data(), 14);
    if (gridCount == 0) {
        err = "load_cloud_file: NanoVDB file without grids";
        return false;
    }
    if (codec != 0) {
        err = "load_cloud_file: compressed NanoVDB files are not supported";
        return false;
    }
    size_t off = 16;
    uint64_t firstSize = 0;
    for (uint16_t g = 0; g < gridCount; g++) {
        if (off + 176 > file.size()) {
            err = "load_cloud_file: truncated NanoVDB metadata";
            return false;
        }
        if (g == 0) firstSize = get<uint64_t>(file.data(), off + 8);
        const uint32_t nameSize = get<uint32_t>(file.data(), off + 136);
        off += 176 + nameSize;
    }
    if (off + firstSize > file.size()) {
        err = "load_cloud_file: truncated NanoVDB grid";
        return false;
    }
    grid.assign(file.begin() + off, file.begin() + off + firstSize);
    return true;
}

std::vector<uint8_t> buildNanoVdbFromDense(const float* values, const int dims[3], double dx, const double p0[3],
                                           const std::string& name, int valueType) {
    return buildNanoVdb(values, dims, dx, p0, name, valueType, DenseBuildOptions());
}

std::vector<uint8_t> buildNanoVdb(const float* values, const int dims[3], double dx, const double p0[3],
                                  const std::string& name, int valueType, const DenseBuildOptions& opt) {
    const bool fp16 = valueType == 1;
    const uint32_t leafSize = fp16 ? kLeafSizeFp16 : kLeafSize;
    // 1. Leaves: every 8^3 block with at least one non-background voxel (GridBuilder.h:251-300); a fully active
    //    constant leaf becomes an active tile of its lower node.
    std::vector<Leaf> leaves;
    std::map<Coord, int> leafIndex;
    const int lx = (dims[0] + 7) / 8, ly = (dims[1] + 7) / 8, lz = (dims[2] + 7) / 8;
    std::vector<std::pair<Coord, float>> leafTiles;
    {
        // blocks scanned in parallel per z slab of blocks, gathered in (bz, by, bx) order
        std::vector<std::vector<Leaf>> slabLeaves(static_cast<size_t>(lz));
        std::vector<std::vector<std::pair<Coord, float>>> slabTiles(static_cast<size_t>(lz));
        parallelFor(size_t(lz), [&](size_t b0, size_t b1, size_t) {
            for (size_t bzi = b0; bzi < b1; bzi++) {
                const int bz = int(bzi);
                for (int by = 0; by < ly; by++)
                    for (int bx = 0; bx < lx; bx++) {
                        Leaf L;
                        L.origin = {opt.ijk0[0] + bx * 8, opt.ijk0[1] + by * 8, opt.ijk0[2] + bz * 8};
                        std::fill(std::begin(L.values), std::end(L.values), 0.0f);
                        std::fill(std::begin(L.mask), std::end(L.mask), 0ull);
                        int active = 0;
                        for (int x = 0; x < 8; x++)
                            for (int y = 0; y < 8; y++)
                                for (int z = 0; z < 8; z++) {
                                    const int gx = bx * 8 + x, gy = by * 8 + y, gz = bz * 8 + z;
                                    if (gx >= dims[0] || gy >= dims[1] || gz >= dims[2]) continue;
                                    const size_t idx = size_t(gx) + (size_t(gy) + size_t(gz) * dims[1]) * dims[0];
                                    const float v = values[idx];
                                    if (opt.active ? !opt.active[idx] : v == 0.0f) continue;
                                    const int n = x << 6 | y << 3 | z;
                                    L.values[n] = v;
                                    L.mask[n >> 6] |= 1ull << (n & 63);
                                    active++;
                                }
                        if (!active) continue;
                        if (active == 512) {
                            bool constant = true;
                            for (int n = 1; n < 512 && constant; n++) constant = L.values[n] == L.values[0];
                            if (constant) {
                                slabTiles[bzi].push_back({L.origin, L.values[0]});
                                continue;
                            }
                        }
                        slabLeaves[bzi].push_back(L);
                    }
            }
        }, 1);
        size_t total = 0;
        for (auto& v : slabLeaves) total += v.size();
        leaves.reserve(total);
        for (size_t z = 0; z < slabLeaves.size(); z++) {
            for (Leaf& L : slabLeaves[z]) {
                leafIndex[L.origin] = int(leaves.size());
                leaves.push_back(L);
            }
            std::vector<Leaf>().swap(slabLeaves[z]);
            for (auto& t : slabTiles[z]) leafTiles.push_back(t);
        }
    }
    // 2. Internal nodes, then the reference's pruning (GridBuilder.h:303-343): a lower node holding only tiles of
    //    one value becomes an active upper tile, an upper node holding only tiles of one value a root tile.
    std::map<Coord, int> lowerIndex, upperIndex;
    std::vector<Lower> lowers;
    std::vector<Upper> uppers;
    auto lowerOf = [&](const Coord& c) -> Lower& {
        const Coord o{c.x & ~127, c.y & ~127, c.z & ~127};
        auto it = lowerIndex.find(o);
        if (it != lowerIndex.end()) return lowers[size_t(it->second)];
        lowerIndex[o] = int(lowers.size());
        lowers.emplace_back();
        Lower& n = lowers.back();
        n.origin = o;
        std::fill(std::begin(n.tileValue), std::end(n.tileValue), 0.0f);
        std::fill(std::begin(n.child), std::end(n.child), -1);
        return n;
    };
    auto lowerSlot = [](const Coord& c) {
        return ((c.x & 127) >> 3) << 8 | ((c.y & 127) >> 3) << 4 | ((c.z & 127) >> 3);
    };
    auto upperSlot = [](const Coord& c) {
        return ((c.x & 4095) >> 7) << 10 | ((c.y & 4095) >> 7) << 5 | ((c.z & 4095) >> 7);
    };
    for (auto& kv : leafIndex) {
        Lower& n = lowerOf(kv.first);
        const int i = lowerSlot(kv.first);
        n.child[i] = kv.second;
        n.childMask[i >> 6] |= 1ull << (i & 63);
    }
    for (auto& t : leafTiles) {
        Lower& n = lowerOf(t.first);
        const int i = lowerSlot(t.first);
        n.tileValue[i] = t.second;
        n.valueMask[i >> 6] |= 1ull << (i & 63);
    }
    auto upperOf = [&](const Coord& c) -> Upper& {
        const Coord o{c.x & ~4095, c.y & ~4095, c.z & ~4095};
        auto it = upperIndex.find(o);
        if (it != upperIndex.end()) return uppers[size_t(it->second)];
        upperIndex[o] = int(uppers.size());
        uppers.emplace_back();
        Upper& n = uppers.back();
        n.origin = o;
        std::fill(std::begin(n.tileValue), std::end(n.tileValue), 0.0f);
        std::fill(std::begin(n.child), std::end(n.child), -1);
        return n;
    };
    for (auto& kv : lowerIndex) {
        Lower& l = lowers[size_t(kv.second)];
        Upper& n = upperOf(kv.first);
        const int i = upperSlot(kv.first);
        bool onlyTiles = true;
        for (int w = 0; w < 64 && onlyTiles; w++) onlyTiles = l.childMask[w] == 0ull;
        bool same = onlyTiles;
        for (int j = 1; j < 4096 && same; j++) same = l.tileValue[j] == l.tileValue[0];
        if (same) {
            n.tileValue[i] = l.tileValue[0];
            n.valueMask[i >> 6] |= 1ull << (i & 63);
            continue;
        }
        n.child[i] = kv.second;
        n.childMask[i >> 6] |= 1ull << (i & 63);
    }
    struct RootTile {
        int upper;  // -1: tile
        float value;
        uint32_t state;
    };
    std::map<Coord, RootTile> rootTable;
    for (auto& kv : upperIndex) {
        const Upper& u = uppers[size_t(kv.second)];
        bool onlyTiles = true;
        for (int w = 0; w < 512 && onlyTiles; w++) onlyTiles = u.childMask[w] == 0ull;
        bool same = onlyTiles;
        for (int j = 1; j < 32768 && same; j++) same = u.tileValue[j] == u.tileValue[0];
        if (same) {
            bool allOn = true;
            for (int w = 0; w < 512 && allOn; w++) allOn = u.valueMask[w] == ~0ull;
            rootTable[kv.first] = RootTile{-1, u.tileValue[0], allOn ? 1u : 0u};
        } else {
            rootTable[kv.first] = RootTile{kv.second, 0.0f, 0u};
        }
    }
    // 3. Breadth-first offsets in root-table / mask order (GridBuilder.h:346-378).
    std::vector<int> upperOrder, lowerOrder, leafOrder;
    for (auto& kv : rootTable) {
        if (kv.second.upper < 0) continue;
        Upper& u = uppers[size_t(kv.second.upper)];
        upperOrder.push_back(kv.second.upper);
        for (int i = 0; i < 32768; i++)
            if (u.child[i] >= 0) {
                lowerOrder.push_back(u.child[i]);
                Lower& l = lowers[size_t(u.child[i])];
                for (int j = 0; j < 4096; j++)
                    if (l.child[j] >= 0) leafOrder.push_back(l.child[j]);
            }
    }
    const uint64_t tiles = rootTable.size();
    const uint64_t rootOff = kGridSize + kTreeSize;
    const uint64_t upperBase = rootOff + kRootSize + kTileSize * tiles;
    const uint64_t lowerBase = upperBase + uint64_t(kUpperSize) * upperOrder.size();
    const uint64_t leafBase = lowerBase + uint64_t(kLowerSize) * lowerOrder.size();
    const uint64_t total = leafBase + uint64_t(leafSize) * leafOrder.size();
    for (size_t i = 0; i < upperOrder.size(); i++) uppers[size_t(upperOrder[i])].offset = upperBase + kUpperSize * i;
    for (size_t i = 0; i < lowerOrder.size(); i++) lowers[size_t(lowerOrder[i])].offset = lowerBase + kLowerSize * i;
    for (size_t i = 0; i < leafOrder.size(); i++) leaves[size_t(leafOrder[i])].offset = leafBase + leafSize * i;

    std::vector<uint8_t> b(total, 0);
    // 4. Leaves: values, masks and the statistics of GridStats::process(Node0) (GridStats.h:685-708).
    std::vector<Stats> leafStats(leaves.size());
    uint64_t voxelCount = 0;
    std::vector<uint64_t> leafVoxels(leaves.size(), 0);
    parallelFor(leaves.size(), [&](size_t l0, size_t l1, size_t) {
    for (size_t li = l0; li < l1; li++) {
        const Leaf& L = leaves[li];
        Stats& s = leafStats[li];
        // Fp16 (GridBuilder.h processLeafs<Fp16>, NanoVDB.h LeafFnBase / LeafData<Fp16>): minimum and quantum over
        // all 512 values, codes uint16(65535 / (max - min) * (v - min) + 0.5) in double, statistics over the decoded
        // values code * quantum + minimum (float), quantised back with + 0.5.
        float minimum = 0.0f, quantum = 0.0f;
        uint16_t code[512];
        float value[512];
        if (fp16) {
            float mn = 3.402823466e38f, mx = -3.402823466e38f;
            for (int n = 0; n < 512; n++) {
                mn = std::min(mn, L.values[n]);
                mx = std::max(mx, L.values[n]);
            }
            minimum = mn;
            quantum = (mx - mn) / 65535.0f;
            const double encode = 65535.0 / double(mx - mn);
            for (int n = 0; n < 512; n++) {
                const float d = L.values[n] - mn;
                code[n] = uint16_t(encode * double(d) + 0.5);
                const float scaled = float(code[n]) * quantum;  // two roundings, as the reference's decode
                value[n] = scaled + minimum;
            }
        } else {
            std::memcpy(value, L.values, sizeof(value));
        }
        for (int n = 0; n < 512; n++)
            if ((L.mask[n >> 6] >> (n & 63)) & 1ull) {
                const int32_t c[3] = {L.origin.x + (n >> 6), L.origin.y + ((n >> 3) & 7), L.origin.z + (n & 7)};
                s.addBox(c, c);
                s.addValue(value[n]);
            }
        leafVoxels[li] = s.count;
        const size_t o = size_t(L.offset);
        putCoord(b, o, s.bmin);
        const uint8_t dif[3] = {uint8_t(s.bmax[0] - s.bmin[0]), uint8_t(s.bmax[1] - s.bmin[1]),
                                uint8_t(s.bmax[2] - s.bmin[2])};
        std::memcpy(b.data() + o + 12, dif, 3);
        b[o + 15] = 2;  // flags: the leaf holds active values
        std::memcpy(b.data() + o + 16, L.mask, 64);
        if (fp16) {
            auto q = [&](float v) {
                const float t = (v - minimum) / quantum;
                return uint16_t(t + 0.5f);
            };
            put<float>(b, o + 80, minimum);
            put<float>(b, o + 84, quantum);
            put<uint16_t>(b, o + 88, q(s.mn));
            put<uint16_t>(b, o + 90, q(s.mx));
            put<uint16_t>(b, o + 92, q(s.avg()));
            const float dq = s.stddev() / quantum;
            put<uint16_t>(b, o + 94, uint16_t(dq + 0.5f));
            std::memcpy(b.data() + o + 96, code, sizeof(code));
        } else {
            put<float>(b, o + 80, s.mn);
            put<float>(b, o + 84, s.mx);
            put<float>(b, o + 88, s.avg());
            put<float>(b, o + 92, s.stddev());
            std::memcpy(b.data() + o + 96, L.values, 2048);
        }
    }
    }, 256);
    for (uint64_t v : leafVoxels) voxelCount += v;
    // 5. Internal nodes (GridStats.h:628-680): active tiles first (value-mask order), then children.
    uint32_t tileCount[3] = {0, 0, 0};
    std::vector<Stats> lowerStats(lowers.size());
    std::vector<uint64_t> lowerTiles(lowers.size(), 0);
    parallelFor(lowers.size(), [&](size_t n0, size_t n1, size_t) {
    for (size_t ni = n0; ni < n1; ni++) {
        const Lower& n = lowers[ni];
        if (n.offset == 0) continue;  // pruned into an upper tile
        Stats& s = lowerStats[ni];
        const size_t o = size_t(n.offset);
        for (int i = 0; i < 4096; i++) {
            if ((n.valueMask[i >> 6] >> (i & 63)) & 1ull) {
                const int32_t lo[3] = {n.origin.x + ((i >> 8) << 3), n.origin.y + (((i >> 4) & 15) << 3),
                                       n.origin.z + ((i & 15) << 3)};
                const int32_t hi[3] = {lo[0] + 7, lo[1] + 7, lo[2] + 7};
                s.addValue(n.tileValue[i], 512);
                s.addBox(lo, hi);
                lowerTiles[ni]++;
            }
        }
        for (int i = 0; i < 4096; i++) {
            if (n.child[i] >= 0) {
                s.merge(leafStats[size_t(n.child[i])]);
                put<int64_t>(b, o + 1088 + 8 * size_t(i), int64_t(leaves[size_t(n.child[i])].offset - n.offset));
            } else {
                put<float>(b, o + 1088 + 8 * size_t(i), n.tileValue[i]);
            }
        }
        putCoord(b, o, s.bmin);
        putCoord(b, o + 12, s.bmax);
        put<uint64_t>(b, o + 24, s.empty() ? 0u : 2u);
        std::memcpy(b.data() + o + 32, n.valueMask, 512);
        std::memcpy(b.data() + o + 544, n.childMask, 512);
        if (!s.empty()) {
            put<float>(b, o + 1056, s.mn);
            put<float>(b, o + 1060, s.mx);
            put<float>(b, o + 1064, s.avg());
            put<float>(b, o + 1068, s.stddev());
        }
    }
    }, 1);
    for (uint64_t t : lowerTiles) {
        tileCount[0] += uint32_t(t);
        voxelCount += 512ull * t;
    }
    std::vector<Stats> upperStats(uppers.size());
    for (size_t ni = 0; ni < uppers.size(); ni++) {
        const Upper& n = uppers[ni];
        if (n.offset == 0) continue;  // pruned into a root tile
        Stats& s = upperStats[ni];
        const size_t o = size_t(n.offset);
        for (int i = 0; i < 32768; i++) {
            if ((n.valueMask[i >> 6] >> (i & 63)) & 1ull) {
                const int32_t lo[3] = {n.origin.x + ((i >> 10) << 7), n.origin.y + (((i >> 5) & 31) << 7),
                                       n.origin.z + ((i & 31) << 7)};
                const int32_t hi[3] = {lo[0] + 127, lo[1] + 127, lo[2] + 127};
                s.addValue(n.tileValue[i], 4096 * 512);
                s.addBox(lo, hi);
                tileCount[1]++;
                voxelCount += 4096ull * 512ull;
            }
        }
        for (int i = 0; i < 32768; i++) {
            if (n.child[i] >= 0) {
                s.merge(lowerStats[size_t(n.child[i])]);
                put<int64_t>(b, o + 8256 + 8 * size_t(i), int64_t(lowers[size_t(n.child[i])].offset - n.offset));
            } else {
                put<float>(b, o + 8256 + 8 * size_t(i), n.tileValue[i]);
            }
        }
        putCoord(b, o, s.bmin);
        putCoord(b, o + 12, s.bmax);
        put<uint64_t>(b, o + 24, s.empty() ? 0u : 2u);
        std::memcpy(b.data() + o + 32, n.valueMask, 4096);
        std::memcpy(b.data() + o + 4128, n.childMask, 4096);
        if (!s.empty()) {
            put<float>(b, o + 8224, s.mn);
            put<float>(b, o + 8228, s.mx);
            put<float>(b, o + 8232, s.avg());
            put<float>(b, o + 8236, s.stddev());
        }
    }
    // 6. Root and its tiles (GridBuilder.h processRoot, GridStats.h:583-624).
    Stats rootStats;
    {
        size_t t = 0;
        for (auto& kv : rootTable) {
            const size_t o = size_t(rootOff + kRootSize + kTileSize * t);
            const Coord& c = kv.first;
            const uint64_t key = uint64_t(uint32_t(c.z) >> 12) | (uint64_t(uint32_t(c.y) >> 12) << 21) |
                                 (uint64_t(uint32_t(c.x) >> 12) << 42);
            put<uint64_t>(b, o, key);
            if (kv.second.upper >= 0) {
                const Upper& u = uppers[size_t(kv.second.upper)];
                rootStats.merge(upperStats[size_t(kv.second.upper)]);
                put<int64_t>(b, o + 8, int64_t(u.offset - rootOff));
                put<uint32_t>(b, o + 16, 0);
                put<float>(b, o + 20, 0.0f);
            } else {
                put<int64_t>(b, o + 8, 0);
                put<uint32_t>(b, o + 16, kv.second.state);
                put<float>(b, o + 20, kv.second.value);
                if (kv.second.state) {
                    const int32_t lo[3] = {c.x, c.y, c.z}, hi[3] = {c.x + 4095, c.y + 4095, c.z + 4095};
                    rootStats.addBox(lo, hi);
                    rootStats.addValue(kv.second.value, 4096ull * 4096ull * 4096ull);
                    tileCount[2]++;
                    voxelCount += 4096ull * 4096ull * 4096ull;
                }
            }
            t++;
        }
    }
    const bool rootEmpty = rootStats.empty();
    if (rootEmpty) {
        const int32_t emptyMin[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, emptyMax[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
        putCoord(b, rootOff, emptyMin);
        putCoord(b, rootOff + 12, emptyMax);
    } else {
        putCoord(b, rootOff, rootStats.bmin);
        putCoord(b, rootOff + 12, rootStats.bmax);
    }
    put<uint32_t>(b, rootOff + 24, uint32_t(tiles));
    put<float>(b, rootOff + 28, opt.background);
    if (tiles == 0) {
        put<float>(b, rootOff + 32, 0.0f);
        put<float>(b, rootOff + 36, 0.0f);
    } else {
        put<float>(b, rootOff + 32, rootStats.count ? rootStats.mn : 0.0f);
        put<float>(b, rootOff + 36, rootStats.count ? rootStats.mx : 0.0f);
        put<float>(b, rootOff + 40, rootStats.avg());
        put<float>(b, rootOff + 44, rootStats.stddev());
    }
    // 7. Tree.
    const size_t tree = kGridSize;
    put<uint64_t>(b, tree + 0, leafOrder.empty() ? 0 : leafBase - tree);
    put<uint64_t>(b, tree + 8, lowerOrder.empty() ? 0 : lowerBase - tree);
    put<uint64_t>(b, tree + 16, upperOrder.empty() ? 0 : upperBase - tree);
    put<uint64_t>(b, tree + 24, rootOff - tree);
    put<uint32_t>(b, tree + 32, uint32_t(leafOrder.size()));
    put<uint32_t>(b, tree + 36, uint32_t(lowerOrder.size()));
    put<uint32_t>(b, tree + 40, uint32_t(upperOrder.size()));
    put<uint32_t>(b, tree + 44, tileCount[0]);
    put<uint32_t>(b, tree + 48, tileCount[1]);
    put<uint32_t>(b, tree + 52, tileCount[2]);
    put<uint64_t>(b, tree + 56, voxelCount);
    // 8. Grid (GridBuilder.h processGrid; GridStats.h:531-570). CloudData passes GridClass::FogVolume in the
    //    oracle slot of getHandle (CloudData.cpp:1398-1400), so its grids keep the builder default class, Unknown.
    put<uint64_t>(b, 0, kMagic);
    put<uint32_t>(b, 16, kVersion);
    put<uint32_t>(b, 20, rootEmpty ? (4u | 8u | 16u | 32u) : (2u | 4u | 8u | 16u | 32u));
    put<uint32_t>(b, 24, 0);
    put<uint32_t>(b, 28, 1);
    put<uint64_t>(b, 32, total);
    std::strncpy(reinterpret_cast<char*>(b.data() + 40), name.c_str(), 255);
    const size_t map = 296;
    const double invdx = 1.0 / dx;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            put<float>(b, map + 4 * (3 * i + j), i == j ? float(dx) : 0.0f);
            put<float>(b, map + 36 + 4 * (3 * i + j), i == j ? float(invdx) : 0.0f);
            put<double>(b, map + 88 + 8 * (3 * i + j), i == j ? dx : 0.0);
            put<double>(b, map + 160 + 8 * (3 * i + j), i == j ? invdx : 0.0);
        }
        put<float>(b, map + 72 + 4 * i, float(p0[i]));
        put<double>(b, map + 232 + 8 * i, p0[i]);
    }
    put<float>(b, map + 84, 1.0f);
    put<double>(b, map + 256, 1.0);
    // World bounding box: the eight corners of the index box [min, max + 1) through the map.
    if (!rootEmpty) {
        double wmin[3] = {1e308, 1e308, 1e308}, wmax[3] = {-1e308, -1e308, -1e308};
        for (int k = 0; k < 8; k++) {
            const double c[3] = {double((k & 1) ? rootStats.bmax[0] + 1 : rootStats.bmin[0]),
                                 double((k & 2) ? rootStats.bmax[1] + 1 : rootStats.bmin[1]),
                                 double((k & 4) ? rootStats.bmax[2] + 1 : rootStats.bmin[2])};
            double w[3];
            applyMap(dx, p0, c, w);
            for (int a = 0; a < 3; a++) {
                wmin[a] = std::min(wmin[a], w[a]);
                wmax[a] = std::max(wmax[a], w[a]);
            }
        }
        for (int a = 0; a < 3; a++) {
            put<double>(b, 560 + 8 * a, wmin[a]);
            put<double>(b, 584 + 8 * a, wmax[a]);
        }
    } else {
        for (int a = 0; a < 3; a++) {
            put<double>(b, 560 + 8 * a, 1.7976931348623157e308);
            put<double>(b, 584 + 8 * a, -1.7976931348623157e308);
        }
    }
    {
        const double one[3] = {1.0, 1.0, 1.0}, zero[3] = {0.0, 0.0, 0.0};
        double m1[3], m0[3];
        applyMap(dx, p0, one, m1);
        applyMap(dx, p0, zero, m0);
        for (int a = 0; a < 3; a++) put<double>(b, 608 + 8 * a, m1[a] - m0[a]);
    }
    put<uint32_t>(b, 632, opt.gridClass);
    put<uint32_t>(b, 636, fp16 ? kGridTypeFp16 : kGridTypeFloat);
    put<int64_t>(b, 640, 0);
    put<uint32_t>(b, 648, 0);
    // 9. Partial checksum (GridChecksum.h:135-150): CRC-32 of grid, tree and root bytes after the checksum field in
    //    the low word, the empty tail CRC in the high word.
    const uint32_t head = crc32(b.data() + 16, size_t(rootOff + kRootSize + kTileSize * tiles) - 16);
    put<uint64_t>(b, 8, (uint64_t(0xFFFFFFFFu) << 32) | head);
    return b;
}

std::vector<uint8_t> buildFogVolumeSphere(float radius, const float center[3], double voxelSize, double halfWidth,
                                          const std::string& name) {
    // initSphere (Primitives.h:494-541): radius and centre in voxel units, float arithmetic as the reference's
    // ValueT = float template; the origin of the index space is world 0.
    const float background = float(halfWidth * voxelSize);
    DenseBuildOptions opt;
    opt.background = background;
    opt.gridClass = 2;  // sdfToFog leaves the grid a FogVolume
    const float r0 = float(double(radius) / voxelSize);
    const float rmax = float(double(r0) + halfWidth);
    float c[3];
    for (int a = 0; a < 3; a++) c[a] = float(double(center[a]) / voxelSize);
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = int(std::floor(c[a] - rmax));
        hi[a] = int(std::ceil(c[a] + rmax));
    }
    const double p0[3] = {0.0, 0.0, 0.0};
    if (r0 < 1.5f) {  // below the Nyquist frequency: an empty grid
        const int dims[3] = {0, 0, 0};
        return buildNanoVdb(nullptr, dims, voxelSize, p0, name, 0, opt);
    }
    // Dense staging over the 8-aligned index box that holds [lo, hi].
    int dims[3];
    for (int a = 0; a < 3; a++) {
        opt.ijk0[a] = lo[a] >= 0 ? lo[a] & ~7 : -((-lo[a] + 7) & ~7);
        dims[a] = ((hi[a] + 1 - opt.ijk0[a]) + 7) & ~7;
    }
    const size_t n = size_t(dims[0]) * dims[1] * dims[2];
    std::vector<float> values(n, 0.0f);
    std::vector<uint8_t> active(n, 0);
    std::vector<uint8_t> band(n, 0);
    auto at = [&](int i, int j, int k) {
        return size_t(i - opt.ijk0[0]) + (size_t(j - opt.ijk0[1]) + size_t(k - opt.ijk0[2]) * dims[1]) * dims[0];
    };
    auto dist = [&](int i, int j, int k) {
        const float x2 = (float(i) - c[0]) * (float(i) - c[0]);
        const float x2y2 = (float(j) - c[1]) * (float(j) - c[1]) + x2;
        return std::sqrt(x2y2 + (float(k) - c[2]) * (float(k) - c[2])) - r0;
    };
    // Narrow band with the reference's leapfrogging along k (voxels it skips lie outside the band).
    for (int i = lo[0]; i <= hi[0]; i++)
        for (int j = lo[1]; j <= hi[1]; j++)
            for (int k = lo[2], m = 1; k <= hi[2]; k += m) {
                m = 1;
                const float v = dist(i, j, k);
                const float d = v < 0.0f ? -v : v;
                if (double(d) < halfWidth) {
                    const size_t q = at(i, j, k);
                    values[q] = float(voxelSize * double(v));  // distance in world units
                    band[q] = 1;
                } else {
                    m += int(std::floor(double(d) - halfWidth));
                }
            }
    // sdfToFog (GridBuilder.h:613-626): v > 0 -> 0 (inactive); else v > -background ? v / -background : 1. Voxels
    // off the band take the inside (-background) or outside (+background) value of the signed flood fill, which for
    // the sphere is the sign of the distance.
    const float dneg = -background, w = 1.0f / dneg;
    for (int k = opt.ijk0[2]; k < opt.ijk0[2] + dims[2]; k++)
        for (int j = opt.ijk0[1]; j < opt.ijk0[1] + dims[1]; j++)
            for (int i = opt.ijk0[0]; i < opt.ijk0[0] + dims[0]; i++) {
                const size_t q = at(i, j, k);
                float v = band[q] ? values[q] : (dist(i, j, k) < 0.0f ? dneg : background);
                if (v > 0.0f) {
                    values[q] = 0.0f;
                    continue;
                }
                values[q] = v > dneg ? v * w : 1.0f;
                active[q] = 1;
            }
    opt.active = active.data();
    return buildNanoVdb(values.data(), dims, voxelSize, p0, name, 0, opt);
}

}  // namespace crvpt
