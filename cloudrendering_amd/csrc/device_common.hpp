// Device-side building blocks of the gfx950 volumetric path tracer: vector math, the Philox4x32-10 counter-based
// RNG, ray/box clipping, Henyey-Greenstein sampling, sky and light evaluation, and the dense/sparse grid lookups.
// Each function names the reference shader code whose behaviour it reproduces.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vpt_types.h"

namespace crvpt {

#define CRVPT_DEV __device__ __forceinline__

// n / d for n < 2^31 with the divisor's host-computed multiplier (Div31).
CRVPT_DEV uint32_t udiv31(uint32_t n, Div31 d) { return uint32_t((uint64_t(n) * d.m) >> d.s); }

constexpr float kPi = 3.14159265359f;
constexpr float kTwoPi = 2.0f * 3.14159265359f;

struct V3 {
    float x, y, z;
};
CRVPT_DEV V3 v3(float a) { return {a, a, a}; }
CRVPT_DEV V3 v3(float a, float b, float c) { return {a, b, c}; }
CRVPT_DEV V3 v3(const float* p) { return {p[0], p[1], p[2]}; }
CRVPT_DEV V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
CRVPT_DEV V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
CRVPT_DEV V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
CRVPT_DEV V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
CRVPT_DEV V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
CRVPT_DEV V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
CRVPT_DEV V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
CRVPT_DEV V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
CRVPT_DEV float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CRVPT_DEV V3 cross(V3 a, V3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
CRVPT_DEV float length(V3 a) { return sqrtf(dot(a, a)); }
CRVPT_DEV V3 normalize(V3 a) { return a * rsqrtf(dot(a, a)); }
CRVPT_DEV float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
CRVPT_DEV V3 vmin(V3 a, float b) { return {fminf(a.x, b), fminf(a.y, b), fminf(a.z, b)}; }
CRVPT_DEV float avg3(V3 a) { return (a.x + a.y + a.z) * (1.0f / 3.0f); }
CRVPT_DEV float lerpf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
CRVPT_DEV V3 lerp3(V3 a, V3 b, float t) { return {lerpf(a.x, b.x, t), lerpf(a.y, b.y, t), lerpf(a.z, b.z, t)}; }
CRVPT_DEV float smoothstepf(float e0, float e1, float x) {
    float t = fminf(fmaxf((x - e0) / (e1 - e0), 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// ------------------------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as 1, 2, 3", SC 2011).
// A path's stream is the counter (draw block, pixel, global frame, seed offset) under a fixed key, so every
// path's draws are independent of which lane or launch runs it.
// ------------------------------------------------------------------------------------------------------------
CRVPT_DEV uint4 philox4x32_10(uint4 c, uint2 k) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        // 32 x 32 -> 64-bit products: one v_mad_u64_u32 each instead of a v_mul_lo_u32 / v_mul_hi_u32 pair
        const uint64_t p0 = uint64_t(0xD2511F53u) * c.x;
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c.z;
        c = make_uint4(uint32_t(p1 >> 32) ^ c.y ^ k.x, uint32_t(p1), uint32_t(p0 >> 32) ^ c.w ^ k.y, uint32_t(p0));
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}

struct Rng {
    uint32_t id0, id1, id2;  // stream identifiers (counter words 1..3)
    uint32_t block;          // counter word 0: index of the next 4-draw block
    uint32_t b0, b1, b2;     // buffered draws not yet used
    uint32_t avail;

    CRVPT_DEV void init(uint32_t a, uint32_t b, uint32_t c) {
        id0 = a; id1 = b; id2 = c;
        block = 0;
        avail = 0;
    }
    CRVPT_DEV uint32_t nextBits() {
        if (avail == 0) {
            uint4 r = philox4x32_10(make_uint4(block, id0, id1, id2), make_uint2(0x6A09E667u, 0x3C6EF372u));
            block++;
            b0 = r.y; b1 = r.z; b2 = r.w;
            avail = 3;
            return r.x;
        }
        uint32_t v = b0;
        b0 = b1; b1 = b2;
        avail--;
        return v;
    }
    // Uniform in [0, 1).
    CRVPT_DEV float uniform() { return float(nextBits() >> 8) * (1.0f / 16777216.0f); }
    // Uniform in (0, 1]: argument of -log for exponential free-path samples.
    CRVPT_DEV float uniformOpen0() { return float((nextBits() >> 8) + 1u) * (1.0f / 16777216.0f); }
    // Exp(1) sample: -log(1 - xi) as in every reference tracker, with xi never reaching 1.
    CRVPT_DEV float exponential() { return -__logf(uniformOpen0()); }
};

// ------------------------------------------------------------------------------------------------------------
// Geometry (VptUtils.glsl:750-801)
// ------------------------------------------------------------------------------------------------------------
CRVPT_DEV void mulMat4(const float* m, float x, float y, float z, float w, float& ox, float& oy, float& oz,
                       float& ow) {
    ox = m[0] * x + m[4] * y + m[8] * z + m[12] * w;
    oy = m[1] * x + m[5] * y + m[9] * z + m[13] * w;
    oz = m[2] * x + m[6] * y + m[10] * z + m[14] * w;
    ow = m[3] * x + m[7] * y + m[11] * z + m[15] * w;
}

// createCameraRay: unproject the near (z = 0) and far (z = 1) NDC points through inverse(P V).
// M: the column-major inverse view-projection (SceneParams::invViewProj or a copy of it).
CRVPT_DEV void createCameraRay(const float* M, float sx, float sy, V3& x, V3& w) {
    float px, py, pz, pw, tx, ty, tz, tw;
    mulMat4(M, sx, sy, 0.0f, 1.0f, px, py, pz, pw);
    mulMat4(M, sx, sy, 1.0f, 1.0f, tx, ty, tz, tw);
    x = v3(px / pw, py / pw, pz / pw);
    w = normalize(v3(tx / tw, ty / tw, tz / tw) - x);
}

// rayBoxIntersect: slab test with near-zero direction components replaced by 1e-6 and tMin clamped to 0.
CRVPT_DEV bool rayBoxIntersect(V3 bMin, V3 bMax, V3 p, V3 d, float& tMin, float& tMax) {
    d.x = fabsf(d.x) <= 0.000001f ? 0.000001f : d.x;
    d.y = fabsf(d.y) <= 0.000001f ? 0.000001f : d.y;
    d.z = fabsf(d.z) <= 0.000001f ? 0.000001f : d.z;
    V3 cmin = (bMin - p) / d;
    V3 cmax = (bMax - p) / d;
    tMin = fmaxf(fmaxf(fminf(cmin.x, cmax.x), fminf(cmin.y, cmax.y)), fminf(cmin.z, cmax.z));
    tMin = fmaxf(0.0f, tMin);
    tMax = fminf(fminf(fmaxf(cmin.x, cmax.x), fmaxf(cmin.y, cmax.y)), fmaxf(cmin.z, cmax.z));
    return !(tMax <= tMin || tMax <= 0.0f);
}

// ------------------------------------------------------------------------------------------------------------
// Phase function (VptUtils.glsl:67-123)
// ------------------------------------------------------------------------------------------------------------
CRVPT_DEV void orthonormalBasis(V3 d, V3& b, V3& t) {
    V3 other = fabsf(d.z) >= 0.9999f ? v3(1.0f, 0.0f, 0.0f) : v3(0.0f, 0.0f, 1.0f);
    b = normalize(cross(other, d));
    t = normalize(cross(d, b));
}

// Henyey-Greenstein pdf of cosTheta from the scene's phase constants (phaseC, set on the host from g).
CRVPT_DEV float hgPdf(const SceneParams& P, float cosTheta) {
    const float denom = P.phaseC[PHASE_ONE_PLUS_G2] - P.phaseC[PHASE_TWO_G] * cosTheta;
    return P.phaseC[PHASE_PDF_SCALE] * __builtin_amdgcn_rcpf(denom * sqrtf(denom));
}

// evaluatePhase
CRVPT_DEV float evalPhase(const SceneParams& P, V3 a, V3 b) { return hgPdf(P, dot(normalize(a), normalize(b))); }

// ------------------------------------------------------------------------------------------------------------
// Sky and lights (VptUtils.glsl:271-453)
// ------------------------------------------------------------------------------------------------------------
CRVPT_DEV float srgbToLinear(float c) { return c <= 0.04045f ? c / 12.92f : __powf((c + 0.055f) / 1.055f, 2.4f); }

CRVPT_DEV V3 envTexel(const SceneParams& P, int x, int y) {
    x = x % P.envW;
    if (x < 0) x += P.envW;
    y = min(max(y, 0), P.envH - 1);
    const float4 t = reinterpret_cast<const float4*>(P.envImage)[size_t(y) * P.envW + x];
    return v3(t.x, t.y, t.z);
}

// sampleSkybox for the equirectangular image, the DEFAULT gradient sky, BLACK and SINGLE_COLOR.
CRVPT_DEV V3 sampleSky(const SceneParams& P, V3 dir) {
    if (P.envKind == ENV_IMAGE) {
        float u = atan2f(dir.z, dir.x) / kTwoPi + 0.5f;
        float v = -asinf(fminf(fmaxf(dir.y, -1.0f), 1.0f)) / kPi + 0.5f;
        float fx = u * float(P.envW) - 0.5f, fy = v * float(P.envH) - 0.5f;
        float x0 = floorf(fx), y0 = floorf(fy);
        float a = fx - x0, b = fy - y0;
        int ix = int(x0), iy = int(y0);
        V3 c = lerp3(lerp3(envTexel(P, ix, iy), envTexel(P, ix + 1, iy), a),
                     lerp3(envTexel(P, ix, iy + 1), envTexel(P, ix + 1, iy + 1), a), b);
        c = v3(fminf(c.x, 100000.0f), fminf(c.y, 100000.0f), fminf(c.z, 100000.0f));
        return v3(P.envFactor) * c;
    }
    if (P.envKind == ENV_DEFAULT) {
        const float dists[5] = {-1.0f, -0.1f, 0.0f, 0.4f, 1.0f};
        const V3 cols[5] = {v3(0.1f, 0.05f, 0.01f), v3(0.01f, 0.05f, 0.2f), v3(0.8f, 0.9f, 1.0f),
                            v3(0.1f, 0.3f, 1.0f), v3(0.01f, 0.1f, 0.7f)};
        V3 c = cols[0];
#pragma unroll
        for (int i = 0; i < 4; i++) c = lerp3(c, cols[i + 1], smoothstepf(dists[i], dists[i + 1], dir.y));
        if (P.useLinearRGB) c = v3(srgbToLinear(c.x), srgbToLinear(c.y), srgbToLinear(c.z));
        return c;
    }
    if (P.envKind == ENV_BLACK) return v3(0.0f);
    return v3(P.sunIntensity);
}

// sampleLight: the Phong sun lobe of the DEFAULT sky, zero otherwise.
CRVPT_DEV V3 sampleSunLobe(const SceneParams& P, V3 dir) {
    if (P.envKind != ENV_DEFAULT) return v3(0.0f);
    const float phongNorm = 11.0f / (2.0f * 3.14159f);
    float c = fmaxf(0.0f, dot(dir, v3(P.sunDirection)));
    float c2 = c * c, c4 = c2 * c2, c8 = c4 * c4;
    return v3(P.sunIntensity) * (c8 * c2 * phongNorm);
}

CRVPT_DEV V3 skyRadiance(const SceneParams& P, V3 dir) { return sampleSky(P, dir) + sampleSunLobe(P, dir); }

CRVPT_DEV V3 lightPositionWorld(const SceneParams& P, const Light& L) {
    V3 p = v3(L.position);
    if (L.lightSpace == 1u || L.lightSpace == 2u) {
        float hom = (L.lightType == 2u || L.lightSpace == 2u) ? 0.0f : 1.0f;
        float ox, oy, oz, ow;
        mulMat4(P.invView, p.x, p.y, p.z, hom, ox, oy, oz, ow);
        p = v3(ox, oy, oz);
    }
    return p;
}

// getHeadlightDirection: direction to light (or the directional light's vector) and its distance.
CRVPT_DEV V3 lightDirection(const SceneParams& P, Rng& rng, V3 pos, int& lightIdx, float& probFactor,
                            float& dist) {
    if (P.numLights > 1) {
        lightIdx = min(int(rng.uniform() * float(P.numLights)), P.numLights - 1);
        probFactor *= float(P.numLights);
    }
    const Light& L = P.lights[lightIdx];
    V3 lp = lightPositionWorld(P, L);
    V3 dir = (L.lightType == 0u || L.lightType == 1u) ? lp - pos : lp;
    dist = length(dir);
    return dir * (1.0f / (dist + 1e-5f));
}

// sampleHeadlight: emitted radiance of light lightIdx towards pos.
CRVPT_DEV V3 lightRadiance(const SceneParams& P, V3 pos, int lightIdx) {
    const Light& L = P.lights[lightIdx];
    V3 lp = lightPositionWorld(P, L);
    float f = L.intensity;
    if ((L.lightType == 0u || L.lightType == 1u) && L.useDistance != 0u) {
        V3 diff = lp - pos;
        f /= fmaxf(dot(diff, diff), 1e-3f);
    }
    if (L.lightType == 1u) {
        float totalWidth = __cosf(L.spotTotalWidth);
        float falloffStart = L.spotTotalWidth > L.spotFalloffStart ? __cosf(L.spotFalloffStart) : totalWidth;
        V3 p = pos - lp;
        V3 z = normalize(v3(L.spotDirection));
        if (L.lightSpace == 1u || L.lightSpace == 2u) {
            float ox, oy, oz, ow;
            mulMat4(P.invView, z.x, z.y, z.z, 0.0f, ox, oy, oz, ow);
            z = v3(ox, oy, oz);
        }
        float cosTheta = dot(z, p) / (length(z) * length(p));
        f *= smoothstepf(totalWidth, falloffStart, cosTheta);
    }
    return v3(L.color) * f;
}

// ------------------------------------------------------------------------------------------------------------
// Grid lookups. Both grids return the normalised density the reference's sampleCloud returns; STOCHASTIC adds a
// uniform sub-voxel jitter (three draws) before the nearest-voxel read (VptUtils.glsl:515-591, 674-684).
// ------------------------------------------------------------------------------------------------------------
// Dense grids need no lookup memo: a voxel is one load.
struct NoCache {
    CRVPT_DEV uint32_t descents() const { return 0u; }
};

// TRI: the trilinear filter is compiled in; the path kernels run TRI = false for the point filters (nearest,
// stochastic), where the eight-tap code would only hold registers. The TRI = false variant also indexes voxels with
// 32-bit arithmetic: the launcher uses it only for grids below 2^32 voxels.
template <bool TRI>
struct DenseAccessT {
    DenseGrid g;
    using Cache = NoCache;
    CRVPT_DEV float voxel(int i, int j, int k) const {
        if ((unsigned)i >= (unsigned)g.dim[0] || (unsigned)j >= (unsigned)g.dim[1] ||
            (unsigned)k >= (unsigned)g.dim[2])
            return g.borderValue;
        if (!TRI) return g.values[(uint32_t(k) * uint32_t(g.dim[1]) + uint32_t(j)) * uint32_t(g.dim[0]) + uint32_t(i)];
        return g.values[(size_t(k) * g.dim[1] + j) * g.dim[0] + i];
    }
    CRVPT_DEV float voxelAt(int i, int j, int k) const { return voxel(i, j, k); }
    CRVPT_DEV float voxelAt(int i, int j, int k, NoCache&) const { return voxel(i, j, k); }
    // u: texel coordinates (grid dim * normalised texture coordinate); (jx, jy, jz) in (-0.5, 0.5) is the
    // stochastic sub-voxel jitter, ignored by the other filters.
    CRVPT_DEV float lookupTex(const SceneParams& P, V3 u, float jx, float jy, float jz) const {
        if (TRI && P.interpolation == INTERP_TRILINEAR) {
            u = u - v3(0.5f);
            const float fx = floorf(u.x), fy = floorf(u.y), fz = floorf(u.z);
            const int i = int(fx), j = int(fy), k = int(fz);
            const float a = u.x - fx, b = u.y - fy, c = u.z - fz;
            const float c00 = lerpf(voxel(i, j, k), voxel(i + 1, j, k), a);
            const float c10 = lerpf(voxel(i, j + 1, k), voxel(i + 1, j + 1, k), a);
            const float c01 = lerpf(voxel(i, j, k + 1), voxel(i + 1, j, k + 1), a);
            const float c11 = lerpf(voxel(i, j + 1, k + 1), voxel(i + 1, j + 1, k + 1), a);
            return lerpf(lerpf(c00, c10, b), lerpf(c01, c11, b), c);
        }
        if (P.interpolation == INTERP_STOCHASTIC) u = u + v3(jx, jy, jz);
        return voxel(int(floorf(u.x)), int(floorf(u.y)), int(floorf(u.z)));
    }
    CRVPT_DEV V3 toTex(const SceneParams& P, V3 pos) const {
        return v3((pos.x - P.boxMin[0]) * P.worldToTex[0], (pos.y - P.boxMin[1]) * P.worldToTex[1],
                  (pos.z - P.boxMin[2]) * P.worldToTex[2]);
    }
    CRVPT_DEV float lookupWorldJ(const SceneParams& P, V3 pos, float jx, float jy, float jz) const {
        return lookupTex(P, toTex(P, pos), jx, jy, jz);
    }
    CRVPT_DEV float lookupWorldJ(const SceneParams& P, V3 pos, float jx, float jy, float jz, NoCache&) const {
        return lookupTex(P, toTex(P, pos), jx, jy, jz);
    }
    CRVPT_DEV float lookupWorld(const SceneParams& P, Rng& rng, V3 pos) const {
        float jx = 0.0f, jy = 0.0f, jz = 0.0f;
        if (P.interpolation == INTERP_STOCHASTIC) {
            jx = rng.uniform() - 0.5f;
            jy = rng.uniform() - 0.5f;
            jz = rng.uniform() - 0.5f;
        }
        return lookupTex(P, toTex(P, pos), jx, jy, jz);
    }
    CRVPT_DEV float lookupWorld(const SceneParams& P, Rng& rng, V3 pos, NoCache&) const {
        return lookupWorld(P, rng, pos);
    }
};
using DenseAccess = DenseAccessT<true>;
using DenseAccessPoint = DenseAccessT<false>;

}  // namespace crvpt
