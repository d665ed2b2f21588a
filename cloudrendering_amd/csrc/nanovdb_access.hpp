// Device read path for NanoVDB 32.3 grids with FP32 values (GridType::Float), the sparse format the reference
// uploads (VolumetricPathTracingPass.cpp:599-674) and reads through PNanoVDB (VptUtils.glsl:496-580), and with
// 16-bit quantised leaves (GridType::Fp16: 1120-byte leaves, minimum f32 at 80, quantum f32 at 84, u16 codes at 96).
//
// Byte layout (offsets relative to each node, FP32 value type; PNanoVDB.glsl:646-1111 lists the same table):
//   grid  672 B: map at 296 (mInvMatF 9 floats at +36, mVecF 3 floats at +72)
//   tree   64 B: root node offset (u64) at 24, relative to the tree
//   root   64 B: background f32 at 28, tile count u32 at 24, tiles of 32 B follow:
//                key u64 at 0, child i64 at 8 (relative to the root, 0 = tile), state u32 at 16, value f32 at 20
//   upper 32^3 children of 128^3 voxels: child mask at 4128, 8-byte table at 8256 (child offset or tile value)
//   lower 16^3 children of 8^3 voxels:   child mask at 544,  8-byte table at 1088
//   leaf  8^3 voxels: f32 values at 96, x-major (x << 6 | y << 3 | z)
// Root keys pack the 4096^3 tile coordinate as (k >> 12) | (j >> 12) << 21 | (i >> 12) << 42.
#pragma once

#include <type_traits>

#include "device_common.hpp"

namespace crvpt {

// Per-lane memo of the last 8^3 block a lookup read, in two registers (the path kernels run at the register cap,
// where every extra live value spills): `key` = the block's position in the grid's flat block index, with bit 31
// set when the block is a tile (no leaf), and `leafOrTile` = the leaf's byte offset / 32, or the tile value's bits. A
// path's consecutive lookups are about a voxel apart, so most hit the memo; a hit returns the value the tree
// descent would. Lookups outside the flat index (or of a grid without one) descend the tree every time.
// (Memoising the block coordinates and the lower node as well measured slower on MI355X: the extra registers cost
// more than the loads they save.)
struct LeafCache {
    static constexpr uint32_t kTileBit = 0x80000000u;
    uint32_t key = 0x40000000u;  // matches no block index (indices are below 2^27)
    uint32_t leafOrTile = 0;  // leaf byte offset / 32 (nodes are 32-byte aligned), or the tile value's bits
    uint32_t misses = 0;  // block lookups and descents (work counter)
    CRVPT_DEV uint32_t descents() const { return misses; }
};

// a * b + c rounded after the product and after the sum (never fused into one FMA).
CRVPT_DEV float mulThenAdd(float a, float b, float c) {
#pragma clang fp contract(off)
    return a * b + c;
}

// WIDE: 64-bit byte offsets, for grids of 4 GiB and more (up to the 128 GiB the block index encodes); the path
// kernels run the 32-bit variant on smaller grids, where the 64-bit arithmetic would cost registers in the hot loop.
// TRI: the trilinear filter is compiled in (see DenseAccessT).
template <bool WIDE, bool TRI = true>
struct NanoAccessT {
    using Off = typename std::conditional<WIDE, uint64_t, uint32_t>::type;
    NanoGrid g;
    using Cache = LeafCache;

    CRVPT_DEV uint32_t u32(Off off) const { return *reinterpret_cast<const uint32_t*>(g.buf + off); }
    CRVPT_DEV float f32(Off off) const { return *reinterpret_cast<const float*>(g.buf + off); }
    CRVPT_DEV uint32_t u16(Off off) const { return *reinterpret_cast<const uint16_t*>(g.buf + off); }
    CRVPT_DEV uint64_t u64(Off off) const {
        const uint2 v = *reinterpret_cast<const uint2*>(g.buf + off);
        return uint64_t(v.x) | (uint64_t(v.y) << 32);
    }

    // Child of lower node `lower` holding (i, j, k): the leaf offset, or 0 with the block's tile value in `tile`.
    CRVPT_DEV Off leafInLower(Off lower, int i, int j, int k, float& tile) const {
        const uint32_t nl = ((uint32_t(i) & 127u) >> 3) << 8 | ((uint32_t(j) & 127u) >> 3) << 4 |
                            ((uint32_t(k) & 127u) >> 3);
        if (((u32(lower + 544u + 4u * (nl >> 5)) >> (nl & 31u)) & 1u) == 0u) {
            tile = f32(lower + 1088u + 8u * nl);
            return 0u;
        }
        return lower + Off(u64(lower + 1088u + 8u * nl));
    }

    // Descends to the 8^3 block holding (i, j, k): returns the leaf offset, or 0 with the tile value in `tile`.
    CRVPT_DEV Off locateBlock(int i, int j, int k, float& tile) const {
        const uint64_t key = uint64_t(uint32_t(k) >> 12) | (uint64_t(uint32_t(j) >> 12) << 21) |
                             (uint64_t(uint32_t(i) >> 12) << 42);
        Off rt = Off(g.rootOffset) + 64u;
        uint32_t t = 0;
        for (; t < g.tileCount; t++, rt += 32u) {
            if (u64(rt) == key) break;
        }
        if (t == g.tileCount) {
            tile = g.background;
            return 0u;
        }
        const uint64_t child = u64(rt + 8u);
        if (child == 0ull) {
            tile = f32(rt + 20u);
            return 0u;
        }
        const Off upper = Off(g.rootOffset) + Off(child);
        const uint32_t nu = ((uint32_t(i) & 4095u) >> 7) << 10 | ((uint32_t(j) & 4095u) >> 7) << 5 |
                            ((uint32_t(k) & 4095u) >> 7);
        if (((u32(upper + 4128u + 4u * (nu >> 5)) >> (nu & 31u)) & 1u) == 0u) {
            tile = f32(upper + 8256u + 8u * nu);
            return 0u;
        }
        const Off lower = upper + Off(u64(upper + 8256u + 8u * nu));
        return leafInLower(lower, i, j, k, tile);
    }

    CRVPT_DEV float value(int i, int j, int k, LeafCache& c) const {
        // block position in the flat index (arithmetic shifts: blocks of negative indices too)
        const uint32_t ti = uint32_t((i >> 3) - g.blockMin[0]), tj = uint32_t((j >> 3) - g.blockMin[1]),
                       tk = uint32_t((k >> 3) - g.blockMin[2]);
        Off leaf;
        if (ti < uint32_t(g.blockDim[0]) && tj < uint32_t(g.blockDim[1]) && tk < uint32_t(g.blockDim[2])) {
            const uint32_t b = (tk * uint32_t(g.blockDim[1]) + tj) * uint32_t(g.blockDim[0]) + ti;
            if ((c.key & ~LeafCache::kTileBit) != b) {
                // one 4-byte load from the flat block index instead of the root / upper / lower chain of six
                // (entries hold leaf offset / 32)
                const uint32_t l = g.blockLeaf[b];
                c.leafOrTile = l ? l : __float_as_uint(g.blockTile[b]);
                c.key = l ? b : (b | LeafCache::kTileBit);
                c.misses++;
            }
            if (c.key & LeafCache::kTileBit) return __uint_as_float(c.leafOrTile);
            leaf = Off(c.leafOrTile) << 5;
        } else {
            float tile = 0.0f;
            leaf = locateBlock(i, j, k, tile);
            c.misses++;
            if (leaf == 0u) return tile;
        }
        const uint32_t n = (uint32_t(i) & 7u) << 6 | (uint32_t(j) & 7u) << 3 | (uint32_t(k) & 7u);
        if (g.fp16) {
            // code * quantum + minimum with two roundings (pnanovdb_leaf_fp_read_float, PNanoVDB.glsl:1322-1336);
            // the leaf's minimum (+80) and quantum (+84) are read next to the code
            const float2 q = *reinterpret_cast<const float2*>(g.buf + leaf + 80u);
            return mulThenAdd(float(u16(leaf + 96u + 2u * n)), q.y, q.x);
        }
        return f32(leaf + 96u + 4u * n);
    }

    // Root tile scan (one tile for every domain inside one 4096^3 tile), then upper, lower and leaf.
    CRVPT_DEV float value(int i, int j, int k) const {
        LeafCache c;
        return value(i, j, k, c);
    }

    // Voxel (i, j, k) of the dense-equivalent texel grid: index coordinate = index bbox min + (i, j, k).
    CRVPT_DEV float voxelAt(int i, int j, int k) const {
        return value(i + g.indexMin[0], j + g.indexMin[1], k + g.indexMin[2]);
    }
    CRVPT_DEV float voxelAt(int i, int j, int k, LeafCache& c) const {
        return value(i + g.indexMin[0], j + g.indexMin[1], k + g.indexMin[2], c);
    }

    // pnanovdb_grid_world_to_indexf in FP32 with the reference's operation order (PNanoVDB.glsl:2008-2018).
    CRVPT_DEV V3 worldToIndex(V3 c) const {
        const float sx = c.x - g.vec[0], sy = c.y - g.vec[1], sz = c.z - g.vec[2];
        return v3(sx * g.invMat[0] + sy * g.invMat[1] + sz * g.invMat[2],
                  sx * g.invMat[3] + sy * g.invMat[4] + sz * g.invMat[5],
                  sx * g.invMat[6] + sy * g.invMat[7] + sz * g.invMat[8]);
    }

    // sampleCloud for NanoVDB (VptUtils.glsl:515-580, 674-684): raw values, no normalisation; (jx, jy, jz) in
    // (-0.5, 0.5) is the stochastic sub-voxel jitter, ignored by the other filters. The box -> grid map
    // (pos - boxMin) / (boxMax - boxMin) * (gridMax - gridMin) + gridMin followed by world_to_indexf is one
    // precomposed affine map (three FMAs for axis-aligned voxels) instead of three divisions and twelve
    // multiply-adds per tap; the index coordinate differs from the step-by-step evaluation only in rounding.
    CRVPT_DEV float lookupWorldJ(const SceneParams& P, V3 pos, float jx, float jy, float jz, LeafCache& cache) const {
        V3 idx;
        if (!WIDE || g.tapDiag) {  // the narrow accessors are only used for grids with a diagonal map
            idx = v3(pos.x * g.tapMat[0] + g.tapOff[0], pos.y * g.tapMat[4] + g.tapOff[1],
                     pos.z * g.tapMat[8] + g.tapOff[2]);
        } else {
            idx = v3(pos.x * g.tapMat[0] + pos.y * g.tapMat[1] + pos.z * g.tapMat[2] + g.tapOff[0],
                     pos.x * g.tapMat[3] + pos.y * g.tapMat[4] + pos.z * g.tapMat[5] + g.tapOff[1],
                     pos.x * g.tapMat[6] + pos.y * g.tapMat[7] + pos.z * g.tapMat[8] + g.tapOff[2]);
        }
        if (TRI && P.interpolation == INTERP_TRILINEAR) {
            idx = idx - v3(0.5f);
            const float fx = floorf(idx.x), fy = floorf(idx.y), fz = floorf(idx.z);
            const int i = int(fx), j = int(fy), k = int(fz);
            const float a = idx.x - fx, b = idx.y - fy, c = idx.z - fz;
            LeafCache& m = cache;
            const float c00 = lerpf(value(i, j, k, m), value(i + 1, j, k, m), a);
            const float c10 = lerpf(value(i, j + 1, k, m), value(i + 1, j + 1, k, m), a);
            const float c01 = lerpf(value(i, j, k + 1, m), value(i + 1, j, k + 1, m), a);
            const float c11 = lerpf(value(i, j + 1, k + 1, m), value(i + 1, j + 1, k + 1, m), a);
            return lerpf(lerpf(c00, c10, b), lerpf(c01, c11, b), c);
        }
        if (P.interpolation == INTERP_STOCHASTIC) idx = idx + v3(jx, jy, jz);
        return value(int(floorf(idx.x)), int(floorf(idx.y)), int(floorf(idx.z)), cache);
    }
    CRVPT_DEV float lookupWorldJ(const SceneParams& P, V3 pos, float jx, float jy, float jz) const {
        LeafCache c;
        return lookupWorldJ(P, pos, jx, jy, jz, c);
    }
    CRVPT_DEV float lookupWorld(const SceneParams& P, Rng& rng, V3 pos, LeafCache& cache) const {
        float jx = 0.0f, jy = 0.0f, jz = 0.0f;
        if (P.interpolation == INTERP_STOCHASTIC) {
            jx = rng.uniform() - 0.5f;
            jy = rng.uniform() - 0.5f;
            jz = rng.uniform() - 0.5f;
        }
        return lookupWorldJ(P, pos, jx, jy, jz, cache);
    }
    CRVPT_DEV float lookupWorld(const SceneParams& P, Rng& rng, V3 pos) const {
        float jx = 0.0f, jy = 0.0f, jz = 0.0f;
        if (P.interpolation == INTERP_STOCHASTIC) {
            jx = rng.uniform() - 0.5f;
            jy = rng.uniform() - 0.5f;
            jz = rng.uniform() - 0.5f;
        }
        return lookupWorldJ(P, pos, jx, jy, jz);
    }
};
using NanoAccess = NanoAccessT<true>;        // any grid (lookups, tables, transmittance, batched queue)
using NanoAccessNarrow = NanoAccessT<false>;  // grids below 4 GiB with axis-aligned voxels (single-frame path kernels)
using NanoAccessNarrowPoint = NanoAccessT<false, false>;  // ... with a point filter

}  // namespace crvpt
