/*
 * CPU oracle for the volumetric path tracing hot path — TEST INFRASTRUCTURE ONLY (see vpt_oracle.h).
 *
 * Part 1 restates, in FP32 and statement by statement, the reference's GLSL compute shaders so that a CPU run makes
 * the same random draws in the same order as the reference's `Clouds.Compute` would:
 *   Data/Shaders/VPT/Clouds.glsl:73-179            pathTraceSample, seeding, running mean
 *   Data/Shaders/VPT/VptUtils.glsl:46-123           hybrid Taus RNG, basis, HG phase sampling / evaluation
 *   Data/Shaders/VPT/VptUtils.glsl:271-453          sky, sRGB, lights (importanceSampleSkybox without env image)
 *   Data/Shaders/VPT/VptUtils.glsl:583-591,674-684  dense sampleCloudRaw / sampleCloud
 *   Data/Shaders/VPT/VptUtils.glsl:750-801          createCameraRay, rayBoxIntersect
 *   Data/Shaders/VPT/DeltaTracking.glsl:167-345     deltaTracking
 *   Data/Shaders/VPT/RatioTracking.glsl:25-96       ratioTracking (albedo-weighted analog variant)
 *   Data/Shaders/VPT/NextEventTracking.glsl:223-394 nextEventTracking
 *   Data/Shaders/VPT/DeltaTracking.glsl:33-164      deltaTrackingSpectral (path-history average probabilities)
 *   Data/Shaders/VPT/NextEventTracking.glsl:30-221  nextEventTrackingSpectral
 *   Data/Shaders/VPT/NeeTransmittance.glsl:4-119,263-382 calculateTransmittance(Distance) (ratio tracking)
 *   Data/Shaders/VPT/DecompositionTracking.glsl:35-123 analogDecompositionTracking (dense)
 *   Data/Shaders/VPT/ResidualRatioTracking.glsl:33-224 residualRatioTracking (dense)
 *   src/PathTracer/SuperVoxelGrid.cpp:40-405        supervoxel tables
 * The equirectangular environment image is sampled uniformly over the sphere (pdf 1/4pi) for next-event estimation:
 * the reference's octahedral importance mips are outside the hot path.
 *
 * Part 2 is an FP64 exact optical-depth integrator used as the closed-form check of transmittance estimators.
 */
#include "vpt_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const float PI = 3.14159265359f;          // VptUtils.glsl:27
const float TWO_PI = 2 * 3.14159265359f;  // VptUtils.glsl:28

struct vec3 {
    float x, y, z;
    vec3() : x(0), y(0), z(0) {}
    vec3(float a) : x(a), y(a), z(a) {}
    vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit vec3(const float* p) : x(p[0]), y(p[1]), z(p[2]) {}
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator-(vec3 a) { return {-a.x, -a.y, -a.z}; }
inline vec3 operator*(vec3 a, vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline vec3 operator/(vec3 a, vec3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3 operator*(float s, vec3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3 operator/(vec3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline vec3& operator+=(vec3& a, vec3 b) { a = a + b; return a; }
inline float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec3 cross(vec3 a, vec3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline float length(vec3 a) { return std::sqrt(dot(a, a)); }
inline vec3 normalize(vec3 a) { return a * (1.0f / length(a)); }
inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
inline vec3 mix3(vec3 a, vec3 b, float t) { return {mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)}; }
inline float clampf(float v, float lo, float hi) { return std::min(std::max(v, lo), hi); }
inline float smoothstepf(float e0, float e1, float x) {
    float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
inline float signf(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }
inline float fractf(float v) { return v - std::floor(v); }

struct vec4 { float x, y, z, w; };
// GLSL mat4 * vec4 with column-major storage m[col*4 + row].
inline vec4 mul(const float* m, vec4 v) {
    vec4 r;
    r.x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w;
    r.y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w;
    r.z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w;
    r.w = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w;
    return r;
}

/* ---------------- RNG: hybrid Taus (VptUtils.glsl:46-65) ---------------- */
struct Rng {
    uint32_t s[4];
    static uint32_t taus(uint32_t z, int S1, int S2, int S3, uint32_t M) {
        uint32_t b = (((z << S1) ^ z) >> S2);
        return ((z & M) << S3) ^ b;
    }
    float random() {
        s[0] = taus(s[0], 13, 19, 12, 4294967294u);
        s[1] = taus(s[1], 2, 25, 4, 4294967288u);
        s[2] = taus(s[2], 3, 11, 17, 4294967280u);
        s[3] = 1664525u * s[3] + 1013904223u;
        return 2.3283064365387e-10f * float(s[0] ^ s[1] ^ s[2] ^ s[3]);
    }
    void init(uint32_t seed) {
        s[0] = s[1] = s[2] = s[3] = seed;
        for (uint32_t i = 0; i < seed % 7u + 2u; i++) random();
    }
};

/* ---------------- Scene view with derived constants ---------------- */
struct Ctx {
    const OracleScene* sc;
    vec3 boxMin, boxMax;
    int dim[3];
    float vmin, vmax;
    // supervoxel tables (RRT and decomposition)
    int svSize = 8;
    int svGrid[3] = {0, 0, 0};
    std::vector<float> rrt;     // (mu_c, mu_r_bar)
    std::vector<float> decomp;  // (min, max)
    std::vector<uint8_t> occ;
    // in-scattering iterations of residualRatioTracking before it stops (ResidualRatioTracking.glsl:203 fixes 10;
    // oracle_trace_rays lowers it to pin the first traversal on its own)
    int rrtMaxIterations = 10;
};
// in-scattering iterations of the calling thread's last residualRatioTracking path (read by oracle_trace_rays)
thread_local int t_lastRrtIterations = 0;

inline float rawVoxel(const Ctx& c, int x, int y, int z) {
    if (x < 0 || y < 0 || z < 0 || x >= c.dim[0] || y >= c.dim[1] || z >= c.dim[2]) return 0.0f;  // border black
    return c.sc->density[size_t(x) + (size_t(y) + size_t(z) * c.dim[1]) * c.dim[0]];
}

/* Dense texture(gridImage, coord).x with NEAREST or LINEAR filtering and CLAMP_TO_BORDER (black). */
inline float textureFetch(const Ctx& c, vec3 coord) {
    if (c.sc->interpolation == OR_TRILINEAR) {
        float u = coord.x * float(c.dim[0]) - 0.5f;
        float v = coord.y * float(c.dim[1]) - 0.5f;
        float w = coord.z * float(c.dim[2]) - 0.5f;
        float fu = std::floor(u), fv = std::floor(v), fw = std::floor(w);
        int i = int(fu), j = int(fv), k = int(fw);
        float a = u - fu, b = v - fv, g = w - fw;
        float c00 = mixf(rawVoxel(c, i, j, k), rawVoxel(c, i + 1, j, k), a);
        float c10 = mixf(rawVoxel(c, i, j + 1, k), rawVoxel(c, i + 1, j + 1, k), a);
        float c01 = mixf(rawVoxel(c, i, j, k + 1), rawVoxel(c, i + 1, j, k + 1), a);
        float c11 = mixf(rawVoxel(c, i, j + 1, k + 1), rawVoxel(c, i + 1, j + 1, k + 1), a);
        return mixf(mixf(c00, c10, b), mixf(c01, c11, b), g);
    }
    int i = int(std::floor(coord.x * float(c.dim[0])));
    int j = int(std::floor(coord.y * float(c.dim[1])));
    int k = int(std::floor(coord.z * float(c.dim[2])));
    return rawVoxel(c, i, j, k);
}

/* sampleCloudRaw + sampleCloud, dense branch (VptUtils.glsl:583-591, 674-684). gridMin = 0, gridMax = 1. */
inline float sampleCloud(const Ctx& c, Rng& rng, vec3 pos) {
    vec3 coord = (pos - c.boxMin) / (c.boxMax - c.boxMin);
    coord = coord * (vec3(1.0f) - vec3(0.0f)) + vec3(0.0f);
    if (c.sc->interpolation == OR_STOCHASTIC) {
        float r0 = rng.random();
        float r1 = rng.random();
        float r2 = rng.random();
        coord += vec3(r0 - 0.5f, r1 - 0.5f, r2 - 0.5f) / vec3(float(c.dim[0]), float(c.dim[1]), float(c.dim[2]));
    }
    return (textureFetch(c, coord) - c.vmin) / (c.vmax - c.vmin);
}

/* ---------------- Sampling helpers (VptUtils.glsl:67-123) ---------------- */
inline void createOrthonormalBasis(vec3 D, vec3& B, vec3& T) {
    vec3 other = std::fabs(D.z) >= 0.9999f ? vec3(1, 0, 0) : vec3(0, 0, 1);
    B = normalize(cross(other, D));
    T = normalize(cross(D, B));
}

inline vec3 randomDirection(Rng& rng, vec3 D) {
    float r1 = rng.random();
    float r2 = rng.random() * 2 - 1;
    float sqrR2 = r2 * r2;
    float two_pi_by_r1 = TWO_PI * r1;
    float sqrt_of_one_minus_sqrR2 = std::sqrt(1.0f - sqrR2);
    float x = std::cos(two_pi_by_r1) * sqrt_of_one_minus_sqrR2;
    float y = std::sin(two_pi_by_r1) * sqrt_of_one_minus_sqrR2;
    float z = r2;
    vec3 t0, t1;
    createOrthonormalBasis(D, t0, t1);
    return t0 * x + t1 * y + D * z;
}

inline float invertcdf(float G, float xi) {
    float t = (1.0f - G * G) / (1.0f - G + 2.0f * G * xi);
    return (0.5f / G) * ((1.0f + G * G) - t * t);
}

inline vec3 importanceSamplePhase(Rng& rng, float G, vec3 D, float& pdf) {
    if (std::fabs(G) < 0.001f) {
        pdf = 1.0f / (4 * PI);
        return randomDirection(rng, -D);
    }
    float phi = rng.random() * 2 * PI;
    float cosTheta = invertcdf(G, rng.random());
    float sinTheta = std::sqrt(std::max(0.0f, 1.0f - cosTheta * cosTheta));
    vec3 t0, t1;
    createOrthonormalBasis(D, t0, t1);
    pdf = 0.25f / PI * (1.0f - G * G) / std::pow((1.0f + G * G) - 2 * G * cosTheta, 1.5f);
    return sinTheta * std::sin(phi) * t0 + sinTheta * std::cos(phi) * t1 + cosTheta * D;
}

inline float evaluatePhase(float G, vec3 a, vec3 b) {
    float cosTheta = dot(normalize(a), normalize(b));
    return 0.25f / PI * (1.0f - G * G) / std::pow((1.0f + G * G) - 2 * G * cosTheta, 1.5f);
}

/* ---------------- Sky and lights (VptUtils.glsl:271-453) ---------------- */
inline float srgbToLinear1(float c) {
    return c <= 0.04045f ? c / 12.92f : std::pow((c + 0.055f) / 1.055f, 2.4f);
}

inline vec3 envImageTexel(const OracleScene* s, int x, int y) {
    x = ((x % s->env_w) + s->env_w) % s->env_w;   // repeat in u
    y = std::min(std::max(y, 0), s->env_h - 1);    // clamp in v
    const float* p = s->env_image + (size_t(y) * s->env_w + x) * 4;
    return vec3(p[0], p[1], p[2]);
}

inline vec3 sampleSkybox(const Ctx& c, vec3 dir) {
    const OracleScene* s = c.sc;
    if (s->env_kind == OR_ENV_IMAGE) {
        float u = std::atan2(dir.z, dir.x) / TWO_PI + 0.5f;
        float v = -std::asin(clampf(dir.y, -1.0f, 1.0f)) / PI + 0.5f;
        float fx = u * float(s->env_w) - 0.5f, fy = v * float(s->env_h) - 0.5f;
        float x0 = std::floor(fx), y0 = std::floor(fy);
        float a = fx - x0, b = fy - y0;
        int ix = int(x0), iy = int(y0);
        vec3 col = mix3(mix3(envImageTexel(s, ix, iy), envImageTexel(s, ix + 1, iy), a),
                        mix3(envImageTexel(s, ix, iy + 1), envImageTexel(s, ix + 1, iy + 1), a), b);
        col = vec3(std::min(col.x, 100000.0f), std::min(col.y, 100000.0f), std::min(col.z, 100000.0f));
        return vec3(s->env_intensity_factor) * col;  // linear-RGB image and linear output
    }
    if (s->env_kind == OR_ENV_DEFAULT) {
        const vec3 BG_COLORS[5] = {vec3(0.1f, 0.05f, 0.01f), vec3(0.01f, 0.05f, 0.2f), vec3(0.8f, 0.9f, 1.0f),
                                   vec3(0.1f, 0.3f, 1.0f), vec3(0.01f, 0.1f, 0.7f)};
        const float BG_DISTS[5] = {-1.0f, -0.1f, 0.0f, 0.4f, 1.0f};
        vec3 col = BG_COLORS[0];
        for (int i = 0; i < 4; i++) {
            col = mix3(col, BG_COLORS[i + 1], smoothstepf(BG_DISTS[i], BG_DISTS[i + 1], dir.y));
        }
        if (s->use_linear_rgb) col = vec3(srgbToLinear1(col.x), srgbToLinear1(col.y), srgbToLinear1(col.z));
        return col;
    }
    if (s->env_kind == OR_ENV_BLACK) return vec3(0.0f);
    return vec3(s->sun_intensity);  // SINGLE_COLOR
}

inline vec3 sampleLight(const Ctx& c, vec3 dir) {
    const OracleScene* s = c.sc;
    if (s->env_kind != OR_ENV_DEFAULT) return vec3(0.0f);
    int N = 10;
    float phongNorm = float(N + 1) / (2 * 3.14159f);
    return vec3(s->sun_intensity) * std::pow(std::max(0.0f, dot(dir, vec3(s->sun_direction))), float(N)) * phongNorm;
}

inline vec3 importanceSampleSkybox(Rng& rng, float& pdf) {
    pdf = 1.0f / (4.0f * PI);
    return randomDirection(rng, vec3(1.0f, 0.0f, 0.0f));
}

inline vec3 lightPositionWorld(const Ctx& c, const OracleLight& L) {
    vec3 p(L.position);
    if (L.light_space == 1 || L.light_space == 2) {
        float hom = (L.light_type == 2 || L.light_space == 2) ? 0.0f : 1.0f;
        vec4 r = mul(c.sc->inverse_view, vec4{p.x, p.y, p.z, hom});
        p = vec3(r.x, r.y, r.z);
    }
    return p;
}

inline vec3 getHeadlightDirection(const Ctx& c, Rng& rng, vec3 pos, uint32_t& lightIdx, float& lpf, float& dist) {
    int n = c.sc->num_lights;
    if (n > 1) {
        lightIdx = std::min(uint32_t(rng.random() * float(n)), uint32_t(n) - 1u);
        lpf *= float(n);
    }
    const OracleLight& L = c.sc->lights[lightIdx];
    vec3 lp = lightPositionWorld(c, L);
    vec3 dir = (L.light_type == 0 || L.light_type == 1) ? lp - pos : lp;
    dist = length(dir);
    return dir / (dist + 1e-5f);
}

inline vec3 sampleHeadlight(const Ctx& c, vec3 pos, uint32_t lightIdx) {
    const OracleLight& L = c.sc->lights[lightIdx];
    vec3 lp = lightPositionWorld(c, L);
    float f = L.intensity;
    if ((L.light_type == 0 || L.light_type == 1) && L.use_distance != 0u) {
        vec3 diff = lp - pos;
        f /= std::max(dot(diff, diff), 1e-3f);
    }
    if (L.light_type == 1) {
        float totalWidth = std::cos(L.spot_total_width);
        float falloffStart = L.spot_total_width > L.spot_falloff_start ? std::cos(L.spot_falloff_start) : totalWidth;
        vec3 p = pos - lp;
        vec3 z = normalize(vec3(L.spot_direction));
        if (L.light_space == 1 || L.light_space == 2) {
            vec4 r = mul(c.sc->inverse_view, vec4{z.x, z.y, z.z, 0.0f});
            z = vec3(r.x, r.y, r.z);
        }
        float cosTheta = dot(z, p) / (length(z) * length(p));
        f *= smoothstepf(totalWidth, falloffStart, cosTheta);
    }
    return f * vec3(L.color);
}

/* ---------------- Geometry (VptUtils.glsl:750-801) ---------------- */
inline void createCameraRay(const Ctx& c, float sx, float sy, vec3& x, vec3& w) {
    vec4 ndcP{sx, sy, 0, 1};
    vec4 ndcT{sx, sy, 1, 1};
    vec4 viewP = mul(c.sc->inverse_view_proj, ndcP);
    viewP.x /= viewP.w; viewP.y /= viewP.w; viewP.z /= viewP.w;
    vec4 viewT = mul(c.sc->inverse_view_proj, ndcT);
    viewT.x /= viewT.w; viewT.y /= viewT.w; viewT.z /= viewT.w;
    x = vec3(viewP.x, viewP.y, viewP.z);
    w = normalize(vec3(viewT.x, viewT.y, viewT.z) - x);
}

inline bool rayBoxIntersect(vec3 bMin, vec3 bMax, vec3 P, vec3 D, float& tMin, float& tMax) {
    D.x = std::fabs(D.x) <= 0.000001f ? 0.000001f : D.x;
    D.y = std::fabs(D.y) <= 0.000001f ? 0.000001f : D.y;
    D.z = std::fabs(D.z) <= 0.000001f ? 0.000001f : D.z;
    vec3 C_Min = (bMin - P) / D;
    vec3 C_Max = (bMax - P) / D;
    tMin = std::max(std::max(std::min(C_Min.x, C_Max.x), std::min(C_Min.y, C_Max.y)), std::min(C_Min.z, C_Max.z));
    tMin = std::max(0.0f, tMin);
    tMax = std::min(std::min(std::max(C_Min.x, C_Max.x), std::max(C_Min.y, C_Max.y)), std::max(C_Min.z, C_Max.z));
    if (tMax <= tMin || tMax <= 0) return false;
    return true;
}

inline float freePath(Rng& rng, float majorant) {
    return -std::log(std::max(0.0000000001f, 1 - rng.random())) / majorant;
}

/* ---------------- NEE transmittance (NeeTransmittance.glsl:4-119, 263-382) ---------------- */
float calculateTransmittanceDistance(const Ctx& c, Rng& rng, vec3 x, vec3 w, float maxDist, bool limited) {
    const OracleScene* s = c.sc;
    float majorant = s->extinction[0];
    float transmittance = 1.0f;
    float tMin, tMax;
    float currDist = 0.0f;
    if (rayBoxIntersect(c.boxMin, c.boxMax, x, w, tMin, tMax)) {
        x += w * tMin;
        currDist += tMin;
        float d = tMax - tMin;
        while (true) {
            float t = freePath(rng, majorant);
            currDist += t;
            if (t > d || (limited && currDist > maxDist)) break;
            vec3 xNew = x + w * t;
            float density = sampleCloud(c, rng, xNew);
            x = xNew;
            float sigma_n = majorant - s->extinction[0] * density;
            float Pn = sigma_n / majorant;
            transmittance *= Pn;
            d -= t;
        }
    }
    return transmittance;
}

/* First scattering / absorption event of a path (Clouds.glsl:91-96 ScatterEvent; feature maps). */
struct ScatterEvent {
    bool hasValue = false;
    vec3 x;
    vec3 w;
    float pdf_w = 0.0f;
    float depth = 0.0f;
    float density = 0.0f;
};
inline void recordFirst(ScatterEvent* fe, vec3 x, vec3 w, float pdf_w, float density, float depth) {
    if (!fe || fe->hasValue) return;
    fe->hasValue = true;
    fe->x = x;
    fe->w = w;
    fe->pdf_w = pdf_w;
    fe->density = density;
    fe->depth = depth;
}

/* ---------------- Trackers ---------------- */
vec3 deltaTracking(const Ctx& c, Rng& rng, vec3 x, vec3 w, ScatterEvent* fe = nullptr) {  // DeltaTracking.glsl:167-345
    const OracleScene* s = c.sc;
    float majorant = s->extinction[0];
    float absorptionAlbedo = 1.0f - s->scattering_albedo[0];
    float scatteringAlbedo = s->scattering_albedo[0];
    float PA = absorptionAlbedo * s->extinction[0];
    float PS = scatteringAlbedo * s->extinction[0];
    int bounces = 0;
    float tMin, tMax;
    if (rayBoxIntersect(c.boxMin, c.boxMax, x, w, tMin, tMax)) {
        x += w * tMin;
        float d = tMax - tMin;
        while (true) {
            float t = freePath(rng, majorant);
            if (t > d) break;
            vec3 xNew = x + w * t;
            float density = sampleCloud(c, rng, xNew);
            x = xNew;
            float sigma_a = PA * density;
            float sigma_n = majorant - s->extinction[0] * density;
            float Pa = sigma_a / majorant;
            float Pn = sigma_n / majorant;
            float xi = rng.random();
            if (xi < Pa) {
                recordFirst(fe, x, vec3(0.0f), 0.0f, density * s->extinction[0], tMax - d + t);
                return vec3(0.0f);
            }
            if (xi < 1 - Pn) {
                if (s->max_bounces > 0 && ++bounces > s->max_bounces) return vec3(0.0f);
                float pdf_w;
                w = importanceSamplePhase(rng, s->phase_g, w, pdf_w);
                recordFirst(fe, x, w, pdf_w, density * s->extinction[0], tMax - d + t);
                if (rayBoxIntersect(c.boxMin, c.boxMax, x, w, tMin, tMax)) {
                    x += w * tMin;
                    d = tMax - tMin;
                }
            } else {
                d -= t;
            }
        }
    }
    (void)PS;
    return sampleSkybox(c, w) + sampleLight(c, w);
}

vec3 ratioTracking(const Ctx& c, Rng& rng, vec3 x, vec3 w, ScatterEvent* fe = nullptr) {  // RatioTracking.glsl:25-96
    const OracleScene* s = c.sc;
    float majorant = s->extinction[0];
    float PA = (1.0f - s->scattering_albedo[0]) * s->extinction[0];
    float PS = s->scattering_albedo[0] * s->extinction[0];
    float transmittance = 1.0f;
    int bounces = 0;
    float tMin, tMax;
    if (rayBoxIntersect(c.boxMin, c.boxMax, x, w, tMin, tMax)) {
        x += w * tMin;
        float d = tMax - tMin;
        while (true) {
            float t = freePath(rng, majorant);
            if (t > d) break;
            x += w * t;
            float density = sampleCloud(c, rng, x);
            float sigma_a = PA * density;
            float sigma_s = PS * density;
            float sigma_n = majorant - s->extinction[0] * density;
            float Pa = sigma_a / majorant;
            float Ps = sigma_s / majorant;
            float Pn = sigma_n / majorant;
            float xi = rng.random();
            if (xi < 1 - Pn) {
                if (s->max_bounces > 0 && ++bounces > s->max_bounces) return vec3(0.0f);
                transmittance *= 1.f - Pa / (Pa + Ps);
                float pdf_w;
                w = importanceSamplePhase(rng, s->phase_g, w, pdf_w);
                recordFirst(fe, x, w, pdf_w, 0.0f, 0.0f);  // ratio tracking sets neither depth nor density
                if (rayBoxIntersect(c.boxMin, c.boxMax, x, w, tMin, tMax)) {
                    x += w * tMin;
                    d = tMax - tMin;
                }
            } else {
                d -= t;
            }
        }
    }
    return transmittance * (sampleSkybox(c, w) + sampleLight(c, w));
}

vec3 nextEventTracking(const Ctx& c, Rng& rng, vec3 x, vec3 w, ScatterEvent* fe = nullptr) {  // NextEventTracking.glsl:223-394
    const OracleScene* s = c.sc;
    float majorant = s->extinction[0];
    float PA = (1.0f - s->scattering_albedo[0]) * s->extinction[0];
    vec3 color(0.0f);
    bool hasValue = false;
    int bounces = 0;
    float tMin, tMax;
    if (rayBoxIntersect(c.boxMin, c.boxMax, x, w, tMin, tMax)) {
        x += w * tMin;
        float d = tMax - tMin;
        while (true) {
            float t = freePath(rng, majorant);
            if (t > d) break;
            vec3 xNew = x + w * t;
            float density = sampleCloud(c, rng, xNew);
            x = xNew;
            float sigma_a = PA * density;
            float sigma_n = majorant - s->extinction[0] * density;
            float Pa = sigma_a / majorant;
            float Pn = sigma_n / majorant;
            float xi = rng.random();
            const float extMax = std::max(std::max(s->extinction[0], s->extinction[1]), s->extinction[2]);
            if (xi < Pa) {
                recordFirst(fe, x, w, 0.0f, density * extMax, tMax - d + t);
                return color;
            }
            if (xi < 1 - Pn) {
                if (s->max_bounces > 0 && ++bounces > s->max_bounces) return color;
                float pdf_w;
                vec3 next_w = importanceSamplePhase(rng, s->phase_g, w, pdf_w);
                recordFirst(fe, x, next_w, pdf_w, density * extMax, tMax - d + t);
                hasValue = true;
                float pdfLightNee = 0.0f;
                vec3 dirLightNee;
                bool useLights = s->num_lights > 0;
                bool isSamplingHeadlight = false;
                float lightProbabilityFactor = 1.0f, lightDistance = 0.0f;
                uint32_t lightIdx = 0;
                if (useLights) {
                    isSamplingHeadlight = (s->is_env_map_black != 0) ? true : (rng.random() > 0.5f);
                    lightProbabilityFactor = s->is_env_map_black != 0 ? 1.0f : 2.0f;
                    if (isSamplingHeadlight) {
                        dirLightNee = getHeadlightDirection(c, rng, x, lightIdx, lightProbabilityFactor, lightDistance);
                    } else {
                        dirLightNee = importanceSampleSkybox(rng, pdfLightNee);
                    }
                } else {
                    dirLightNee = importanceSampleSkybox(rng, pdfLightNee);
                }
                float pdf_nee_phase = evaluatePhase(s->phase_g, w, dirLightNee);
                w = next_w;
                vec3 colorNew;
                if (useLights) {
                    float commonFactor = lightProbabilityFactor * pdf_nee_phase;
                    if (isSamplingHeadlight) {
                        colorNew = (commonFactor * calculateTransmittanceDistance(c, rng, x, dirLightNee,
                                                                                  lightDistance, true)) *
                                   sampleHeadlight(c, x, lightIdx);
                    } else {
                        colorNew = (commonFactor / pdfLightNee *
                                    calculateTransmittanceDistance(c, rng, x, dirLightNee, 0.0f, false)) *
                                   (sampleSkybox(c, dirLightNee) + sampleLight(c, dirLightNee));
                    }
                } else {
                    colorNew = (pdf_nee_phase / pdfLightNee *
                                calculateTransmittanceDistance(c, rng, x, dirLightNee, 0.0f, false)) *
                               (sampleSkybox(c, dirLightNee) + sampleLight(c, dirLightNee));
                }
                color += colorNew;
                if (rayBoxIntersect(c.boxMin, c.boxMax, x, w, tMin, tMax)) {
                    x += w * tMin;
                    d = tMax - tMin;
                }
            } else {
                d -= t;
            }
        }
    }
    if (!hasValue) color += sampleSkybox(c, w) + sampleLight(c, w);
    return color;
}

/* ---------------- Spectral trackers (Kutz et al. 2017; path-history average probabilities) ---------------- */
inline float maxComponent(vec3 v) { return std::max(std::max(v.x, v.y), v.z); }
inline float avgComponent(vec3 v) { return (v.x + v.y + v.z) / 3.0f; }
inline vec3 min3(vec3 a, float b) { return {std::min(a.x, b), std::min(a.y, b), std::min(a.z, b)}; }

vec3 deltaTrackingSpectral(const Ctx& c, Rng& rng, vec3 x, vec3 w, ScatterEvent* fe = nullptr) {  // DeltaTracking.glsl:33-164
    const OracleScene* s = c.sc;
    const vec3 extinction(s->extinction), albedo(s->scattering_albedo);
    float majorant = maxComponent(extinction);
    vec3 weights(1.0f);
    vec3 absorptionAlbedo = vec3(1.0f) - albedo;
    vec3 scatteringAlbedo = albedo;
    int bounces = 0;
    float tMin, tMax;
    if (rayBoxIntersect(c.boxMin, c.boxMax, x, w, tMin, tMax)) {
        x += w * tMin;
        float d = tMax - tMin;
        while (true) {
            float t = freePath(rng, majorant);
            if (t > d) break;
            vec3 xNew = x + w * t;
            float density = sampleCloud(c, rng, xNew);
            x = xNew;
            vec3 sigma_a = absorptionAlbedo * extinction * density;
            vec3 sigma_s = scatteringAlbedo * extinction * density;
            vec3 sigma_n = vec3(majorant) - extinction * density;
            float Pa = avgComponent(sigma_a * weights);
            float Ps = avgComponent(sigma_s * weights);
            float Pn = avgComponent(sigma_n * weights);
            float C = Pa + Ps + Pn;
            Pa /= C;
            Ps /= C;
            Pn /= C;
            float xi = rng.random();
            if (xi < Pa) {
                recordFirst(fe, x, vec3(0.0f), 0.0f, density * majorant, tMax - d + t);
                return vec3(0.0f);
            }
            if (xi < Pa + Ps) {
                if (s->max_bounces > 0 && ++bounces > s->max_bounces) return vec3(0.0f);
                float pdf_w;
                w = importanceSamplePhase(rng, s->phase_g, w, pdf_w);
                recordFirst(fe, x, w, pdf_w, density * majorant, tMax - d + t);
                if (rayBoxIntersect(c.boxMin, c.boxMax, x, w, tMin, tMax)) {
                    x += w * tMin;
                    d = tMax - tMin;
                }
                weights = weights * (sigma_s / (majorant * Ps));
            } else {
                d -= t;
                weights = weights * (sigma_n / (majorant * Pn));
            }
            weights = min3(weights, 100.0f);
        }
    }
    return min3(weights, 100000.0f) * (sampleSkybox(c, w) + sampleLight(c, w));
}

vec3 nextEventTrackingSpectral(const Ctx& c, Rng& rng, vec3 x, vec3 w, ScatterEvent* fe = nullptr) {  // NextEventTracking.glsl:30-221
    const OracleScene* s = c.sc;
    const vec3 extinction(s->extinction), albedo(s->scattering_albedo);
    float majorant = maxComponent(extinction);
    vec3 weights(1.0f);
    vec3 absorptionAlbedo = vec3(1.0f) - albedo;
    vec3 scatteringAlbedo = albedo;
    vec3 color(0.0f);
    bool hasValue = false;
    int bounces = 0;
    float tMin, tMax;
    if (rayBoxIntersect(c.boxMin, c.boxMax, x, w, tMin, tMax)) {
        x += w * tMin;
        float d = tMax - tMin;
        while (true) {
            float t = freePath(rng, majorant);
            if (t > d) break;
            vec3 xNew = x + w * t;
            float density = sampleCloud(c, rng, xNew);
            x = xNew;
            vec3 sigma_a = absorptionAlbedo * extinction * density;
            vec3 sigma_s = scatteringAlbedo * extinction * density;
            vec3 sigma_n = vec3(majorant) - extinction * density;
            float Pa = avgComponent(sigma_a * weights);
            float Ps = avgComponent(sigma_s * weights);
            float Pn = avgComponent(sigma_n * weights);
            float C = Pa + Ps + Pn;
            Pa /= C;
            Ps /= C;
            Pn /= C;
            float xi = rng.random();
            if (xi < Pa) {
                recordFirst(fe, x, vec3(0.0f), 0.0f, density * majorant, tMax - d + t);
                return color;
            }
            if (xi < Pa + Ps) {
                if (s->max_bounces > 0 && ++bounces > s->max_bounces) return color;
                float pdf_w;
                vec3 next_w = importanceSamplePhase(rng, s->phase_g, w, pdf_w);
                recordFirst(fe, x, next_w, pdf_w, density * majorant, tMax - d + t);
                hasValue = true;
                float pdfLightNee = 0.0f;
                vec3 dirLightNee;
                bool useLights = s->num_lights > 0;
                bool isSamplingHeadlight = false;
                float lightProbabilityFactor = 1.0f, lightDistance = 0.0f;
                uint32_t lightIdx = 0;
                if (useLights) {
                    isSamplingHeadlight = (s->is_env_map_black != 0) ? true : (rng.random() > 0.5f);
                    lightProbabilityFactor = s->is_env_map_black != 0 ? 1.0f : 2.0f;
                    if (isSamplingHeadlight) {
                        dirLightNee = getHeadlightDirection(c, rng, x, lightIdx, lightProbabilityFactor, lightDistance);
                    } else {
                        dirLightNee = importanceSampleSkybox(rng, pdfLightNee);
                    }
                } else {
                    dirLightNee = importanceSampleSkybox(rng, pdfLightNee);
                }
                float pdf_nee_phase = evaluatePhase(s->phase_g, w, dirLightNee);
                w = next_w;
                weights = weights * (sigma_s / (majorant * Ps));
                if (useLights) {
                    vec3 commonFactor = (lightProbabilityFactor * pdf_nee_phase) * min3(weights, 100000.0f);
                    if (isSamplingHeadlight) {
                        color += commonFactor *
                                 calculateTransmittanceDistance(c, rng, x, dirLightNee, lightDistance, true) *
                                 sampleHeadlight(c, x, lightIdx);
                    } else {
                        color += commonFactor / pdfLightNee *
                                 calculateTransmittanceDistance(c, rng, x, dirLightNee, 0.0f, false) *
                                 (sampleSkybox(c, dirLightNee) + sampleLight(c, dirLightNee));
                    }
                } else {
                    color += (pdf_nee_phase / pdfLightNee *
                              calculateTransmittanceDistance(c, rng, x, dirLightNee, 0.0f, false)) *
                             min3(weights, 100000.0f) * (sampleSkybox(c, dirLightNee) + sampleLight(c, dirLightNee));
                }
                if (rayBoxIntersect(c.boxMin, c.boxMax, x, w, tMin, tMax)) {
                    x += w * tMin;
                    d = tMax - tMin;
                }
            } else {
                d -= t;
                weights = weights * (sigma_n / (majorant * Pn));
            }
            weights = min3(weights, 100.0f);
        }
    }
    if (!hasValue) color += min3(weights, 100000.0f) * (sampleSkybox(c, w) + sampleLight(c, w));
    return color;
}

inline int svIndex(const Ctx& c, int x, int y, int z) { return x + (y + z * c.svGrid[1]) * c.svGrid[0]; }
inline bool svInside(const Ctx& c, int x, int y, int z) {
    return x >= 0 && y >= 0 && z >= 0 && x < c.svGrid[0] && y < c.svGrid[1] && z < c.svGrid[2];
}

vec3 analogDecompositionTracking(const Ctx& c, Rng& rng, vec3 x, vec3 w) {  // DecompositionTracking.glsl:35-123
    const OracleScene* s = c.sc;
    const vec3 EPSILON_VEC(1e-6f);
    float tMinVal, tMaxVal;
    if (rayBoxIntersect(c.boxMin + EPSILON_VEC, c.boxMax - EPSILON_VEC, x, w, tMinVal, tMaxVal)) {
        float majorant = s->extinction[0];
        float absorptionAlbedo = 1.0f - s->scattering_albedo[0];
        vec3 voxelGridSize(float(c.dim[0]), float(c.dim[1]), float(c.dim[2]));
        vec3 boxDelta = c.boxMax - c.boxMin;
        vec3 superVoxelSize = vec3(float(c.svSize)) * boxDelta / voxelGridSize;
        x += w * tMinVal;
        vec3 startPoint = (x - c.boxMin) / boxDelta * voxelGridSize / vec3(float(c.svSize));
        int idx[3] = {int(std::floor(startPoint.x)), int(std::floor(startPoint.y)), int(std::floor(startPoint.z))};
        int cached[3] = {-1, -1, -1};
        int bounces = 0;
        float mmMin = 0.0f, mmMax = 0.0f;
        while (svInside(c, idx[0], idx[1], idx[2])) {
            vec3 minP = c.boxMin + vec3(float(idx[0]), float(idx[1]), float(idx[2])) * superVoxelSize;
            vec3 maxP = minP + superVoxelSize;
            float tMinSV = 0.0f, tMaxSV = 0.0f;
            rayBoxIntersect(minP, maxP, x, w, tMinSV, tMaxSV);
            float d_max = tMaxSV - tMinSV;
            x += w * tMinSV;
            if (cached[0] != idx[0] || cached[1] != idx[1] || cached[2] != idx[2]) {
                int si = svIndex(c, idx[0], idx[1], idx[2]);
                mmMin = c.decomp[2 * si];
                mmMax = c.decomp[2 * si + 1];
                cached[0] = idx[0]; cached[1] = idx[1]; cached[2] = idx[2];
            }
            if (mmMax < 1e-5f) {
                x += w * d_max;
            } else {
                float mu_c_t = std::max(0.0000000001f, majorant * mmMin);
                float majorant_r_local = std::max(0.0000000001f, majorant * mmMax - mu_c_t);
                float t_c = freePath(rng, mu_c_t);
                float t_r = 0.0f;
                while (true) {
                    t_r -= std::log(std::max(0.0000000001f, 1 - rng.random())) / majorant_r_local;
                    if (t_c >= d_max && t_r >= d_max) {
                        x = x + d_max * w;
                        break;
                    }
                    vec3 xs = x + w * std::min(t_c, t_r);
                    bool isCollision = false;
                    if (t_c <= t_r) {
                        isCollision = true;
                    } else {
                        float density = sampleCloud(c, rng, xs);
                        isCollision = rng.random() * majorant_r_local < s->extinction[0] * density - mu_c_t;
                    }
                    if (isCollision) {
                        x = xs;
                        if (rng.random() < absorptionAlbedo) return vec3(0.0f);
                        if (s->max_bounces > 0 && ++bounces > s->max_bounces) return vec3(0.0f);
                        float pdf_w;
                        w = importanceSamplePhase(rng, s->phase_g, w, pdf_w);
                        t_r = 0.0f;
                        t_c = freePath(rng, mu_c_t);
                        rayBoxIntersect(minP, maxP, x, w, tMinSV, tMaxSV);
                        d_max = tMaxSV - tMinSV;
                    }
                }
            }
            vec3 cellCenter = (minP + maxP) * 0.5f;
            vec3 mov = x + w * 0.00001f - cellCenter;
            vec3 smov(signf(mov.x), signf(mov.y), signf(mov.z));
            mov = mov * smov;
            int dx = (mov.x >= mov.y && mov.x >= mov.z) ? 1 : 0;
            int dy = (mov.y >= mov.x && mov.y >= mov.z) ? 1 : 0;
            int dz = (mov.z >= mov.x && mov.z >= mov.y) ? 1 : 0;
            idx[0] += dx * int(smov.x);
            idx[1] += dy * int(smov.y);
            idx[2] += dz * int(smov.z);
        }
    }
    return sampleSkybox(c, w) + sampleLight(c, w);
}

float residualRatioTrackingEstimator(const Ctx& c, Rng& rng, vec3 x, vec3 w, float dStart, float dEnd, float T,
                                     float& reservoirWeightSum, float& reservoirT, float& reservoirDist,
                                     float mu_c, float mu_r_bar) {  // ResidualRatioTracking.glsl:33-76
    float T_c = std::exp(-mu_c * (dEnd - dStart));
    float T_r = 1.0f;
    float dTravelled = dStart;
    do {
        float t = freePath(rng, mu_r_bar);
        x += w * t;
        dTravelled += t;
        if (dTravelled >= dEnd) break;
        float density = sampleCloud(c, rng, x);
        float mu = c.sc->extinction[0] * density;
        T_r *= 1.0f - (mu - mu_c) / mu_r_bar;
        float T_c_local = std::exp(-mu_c * (dTravelled - dStart));
        float T_local = T * T_r * T_c_local;
        float reservoirWeight = T_local;
        reservoirWeightSum += reservoirWeight;
        float xi = rng.random();
        if (xi < reservoirWeight / reservoirWeightSum) {
            reservoirT = T_local;
            reservoirDist = dTravelled;
        }
    } while (true);
    return T_c * T_r;
}

/* Supervoxel DDA set-up shared by the RRT renderer and the RRT transmittance estimator
 * (ResidualRatioTracking.glsl:104-156). Returns false when all steps are zero. */
struct SvDda {
    int step[3];
    float tMaxA[3], tDeltaA[3];
    int idx[3];
};
inline bool svDdaInit(const Ctx& c, vec3 x, vec3 w, float dTotal, SvDda& q) {
    vec3 voxelGridSize(float(c.dim[0]), float(c.dim[1]), float(c.dim[2]));
    vec3 boxDelta = c.boxMax - c.boxMin;
    vec3 startPoint = (x - c.boxMin) / boxDelta * voxelGridSize / vec3(float(c.svSize));
    vec3 endPoint = (x + w * dTotal - c.boxMin) / boxDelta * voxelGridSize / vec3(float(c.svSize));
    for (int a = 0; a < 3; a++) {
        int stepA = int(signf(endPoint[a] - startPoint[a]));
        float tDelta;
        if (stepA != 0) tDelta = std::min(float(stepA) / (endPoint[a] - startPoint[a]), 1e7f);
        else tDelta = 1e7f;
        float tMaxV;
        if (stepA > 0) tMaxV = tDelta * (1.0f - fractf(startPoint[a]));
        else tMaxV = tDelta * fractf(startPoint[a]);
        q.idx[a] = int(std::floor(startPoint[a]));
        q.step[a] = stepA;
        q.tDeltaA[a] = tDelta;
        q.tMaxA[a] = tMaxV;
    }
    return !(q.step[0] == 0 && q.step[1] == 0 && q.step[2] == 0);
}
inline void svDdaAdvance(SvDda& q) {
    if (q.tMaxA[0] < q.tMaxA[1]) {
        if (q.tMaxA[0] < q.tMaxA[2]) { q.idx[0] += q.step[0]; q.tMaxA[0] += q.tDeltaA[0]; }
        else { q.idx[2] += q.step[2]; q.tMaxA[2] += q.tDeltaA[2]; }
    } else {
        if (q.tMaxA[1] < q.tMaxA[2]) { q.idx[1] += q.step[1]; q.tMaxA[1] += q.tDeltaA[1]; }
        else { q.idx[2] += q.step[2]; q.tMaxA[2] += q.tDeltaA[2]; }
    }
}
inline void svBounds(const Ctx& c, const int idx[3], vec3& mn, vec3& mx) {
    vec3 voxelGridSize(float(c.dim[0]), float(c.dim[1]), float(c.dim[2]));
    vec3 boxDelta = c.boxMax - c.boxMin;
    vec3 minVoxelPos(float(idx[0] * c.svSize), float(idx[1] * c.svSize), float(idx[2] * c.svSize));
    vec3 maxVoxelPos = minVoxelPos + vec3(float(c.svSize));
    mn = minVoxelPos / voxelGridSize * boxDelta + c.boxMin;
    mx = maxVoxelPos / voxelGridSize * boxDelta + c.boxMin;
}

vec3 residualRatioTracking(const Ctx& c, Rng& rng, vec3 x, vec3 w) {  // ResidualRatioTracking.glsl:78-224
    const OracleScene* s = c.sc;
    float reservoirWeightSum = 0.0f, reservoirT = 0.0f, reservoirDist = 0.0f;
    float T = 1.0f;
    vec3 accumulatedColor(0.0f);
    const vec3 EPSILON_VEC(1e-6f);
    float tMinVal, tMaxVal;
    vec3 oldX;
    int iteration = 0;
    while (true) {
        if (rayBoxIntersect(c.boxMin + EPSILON_VEC, c.boxMax - EPSILON_VEC, x, w, tMinVal, tMaxVal)) {
            x += w * tMinVal;
            oldX = x;
            float dTotal = tMaxVal - tMinVal;
            SvDda q;
            if (!svDdaInit(c, x, w, dTotal, q)) break;
            while (svInside(c, q.idx[0], q.idx[1], q.idx[2])) {
                int si = svIndex(c, q.idx[0], q.idx[1], q.idx[2]);
                float mu_c = c.rrt[2 * si];
                float mu_r_bar = c.rrt[2 * si + 1];
                vec3 mn, mx;
                svBounds(c, q.idx, mn, mx);
                float tMinVoxel = 0.0f, tMaxVoxel = 0.0f;
                rayBoxIntersect(mn, mx, oldX, w, tMinVoxel, tMaxVoxel);
                x = oldX + w * tMinVoxel;
                T *= residualRatioTrackingEstimator(c, rng, x, w, tMinVoxel, tMaxVoxel, T, reservoirWeightSum,
                                                    reservoirT, reservoirDist, mu_c, mu_r_bar);
                svDdaAdvance(q);
            }
        } else {
            break;
        }
        float xi = rng.random();
        if (xi > reservoirWeightSum || iteration >= c.rrtMaxIterations) break;
        accumulatedColor += T * (sampleSkybox(c, w) + sampleLight(c, w));
        iteration++;
        T = reservoirT;
        float pdf_w;
        x = oldX + w * reservoirDist;
        w = importanceSamplePhase(rng, s->phase_g, w, pdf_w);
        reservoirWeightSum = 0.0f;
        reservoirT = 0.0f;
        reservoirDist = 0.0f;
    }
    accumulatedColor += T * (sampleSkybox(c, w) + sampleLight(c, w));
    t_lastRrtIterations = iteration;
    return accumulatedColor;
}

/* ---------------- Supervoxel tables (SuperVoxelGrid.cpp:40-405) ---------------- */
inline float normVoxel(const Ctx& c, int x, int y, int z) {
    // DensityField::getDataFloatAtNorm (DensityField.cpp:262-267): (v - min) / (max - min)
    return (c.sc->density[size_t(x) + (size_t(y) + size_t(z) * c.dim[1]) * c.dim[0]] - c.vmin) / (c.vmax - c.vmin);
}

void buildSupervoxels(Ctx& c, float extinction) {
    int sv = std::max(c.sc->super_voxel_size > 0 ? c.sc->super_voxel_size : 8, 1);
    if (c.dim[0] < sv || c.dim[1] < sv || c.dim[2] < sv) {
        bool divisible;
        do {
            divisible = c.dim[0] % sv == 0 && c.dim[1] % sv == 0 && c.dim[2] % sv == 0;
            if (!divisible) sv /= 2;
        } while (!divisible);
    }
    c.svSize = sv;
    for (int a = 0; a < 3; a++) c.svGrid[a] = (c.dim[a] + sv - 1) / sv;
    int n = c.svGrid[0] * c.svGrid[1] * c.svGrid[2];
    c.rrt.assign(2 * size_t(n), 0.0f);
    c.decomp.assign(2 * size_t(n), 0.0f);
    c.occ.assign(size_t(n), 0);
    bool nearest = c.sc->interpolation == OR_NEAREST;
    const float gamma = 2.0f;
    const float D = std::sqrt(3.0f) * float(sv);
    for (int si = 0; si < n; si++) {
        int sx = si % c.svGrid[0], sy = (si / c.svGrid[0]) % c.svGrid[1], sz = si / (c.svGrid[0] * c.svGrid[1]);
        float dmin = std::numeric_limits<float>::max();
        float dmax = std::numeric_limits<float>::lowest();
        float davg = 0.0f;
        int count = 0;
        int lo = nearest ? 0 : -1, hi = nearest ? sv - 1 : sv;
        for (int oz = lo; oz <= hi; oz++)
            for (int oy = lo; oy <= hi; oy++)
                for (int ox = lo; ox <= hi; ox++) {
                    int vx = sx * sv + ox, vy = sy * sv + oy, vz = sz * sv + oz;
                    float value;
                    if (vx >= 0 && vy >= 0 && vz >= 0 && vx < c.dim[0] && vy < c.dim[1] && vz < c.dim[2]) {
                        value = normVoxel(c, vx, vy, vz);
                    } else {
                        value = 0.0f;
                    }
                    dmin = std::min(dmin, value);
                    dmax = std::max(dmax, value);
                    davg += value;
                    count++;
                }
        davg /= float(count);
        c.decomp[2 * si] = dmin;
        c.decomp[2 * si + 1] = dmax;
        if ((c.dim[0] % sv != 0 && sx == c.svGrid[0] - 1) || (c.dim[1] % sv != 0 && sy == c.svGrid[1] - 1) ||
            (c.dim[2] % sv != 0 && sz == c.svGrid[2] - 1)) {
            dmin = 0.0f;  // only the RRT table forces the boundary minimum to zero (SuperVoxelGrid.cpp:180-191)
        }
        float mu_min = dmin * extinction, mu_max = dmax * extinction, mu_avg = davg * extinction;
        float mu_r_bar = std::max(mu_max - mu_min, 0.1f);
        float mu_c = mu_min + mu_r_bar * std::pow(gamma, (1.0f / (D * mu_r_bar)) - 1.0f);
        float mu_c_prime = clampf(mu_c, mu_min, mu_avg);
        c.rrt[2 * si] = mu_c_prime;
        c.rrt[2 * si + 1] = mu_r_bar;
        c.occ[si] = dmax < 1e-5f ? 0 : 1;
    }
}

void initCtx(Ctx& c, const OracleScene* s) {
    c.sc = s;
    c.boxMin = vec3(s->box_min);
    c.boxMax = vec3(s->box_max);
    c.dim[0] = s->nx; c.dim[1] = s->ny; c.dim[2] = s->nz;
    c.vmin = s->voxel_value_min;
    c.vmax = s->voxel_value_max;
    if (c.vmin == c.vmax) c.vmin = 0.0f;  // VolumetricPathTracingPass.cpp:2055-2057
    if (s->mode == OR_RESIDUAL_RATIO || s->mode == OR_DECOMPOSITION) buildSupervoxels(c, s->extinction[0]);
}

vec3 tracePath(const Ctx& c, Rng& rng, vec3 x, vec3 w, ScatterEvent* fe = nullptr) {
    switch (c.sc->mode) {
        case OR_RATIO: return ratioTracking(c, rng, x, w, fe);
        case OR_DECOMPOSITION: return analogDecompositionTracking(c, rng, x, w);  // records no first event
        case OR_RESIDUAL_RATIO: return residualRatioTracking(c, rng, x, w);     // records no first event
        case OR_NEXT_EVENT: return nextEventTracking(c, rng, x, w, fe);
        case OR_DELTA_SPECTRAL: return deltaTrackingSpectral(c, rng, x, w, fe);
        case OR_NEXT_EVENT_SPECTRAL: return nextEventTrackingSpectral(c, rng, x, w, fe);
        default: return deltaTracking(c, rng, x, w, fe);
    }
}

/* ---------------- Transmittance estimators for ray queries ---------------- */
float transmittanceDelta(const Ctx& c, Rng& rng, vec3 x, vec3 w, float maxDist) {
    // Free-flight part of deltaTracking (DeltaTracking.glsl:208-271): any real collision ends the ray.
    float majorant = c.sc->extinction[0];
    float tMin, tMax, currDist = 0.0f;
    if (rayBoxIntersect(c.boxMin, c.boxMax, x, w, tMin, tMax)) {
        x += w * tMin;
        currDist += tMin;
        float d = tMax - tMin;
        while (true) {
            float t = freePath(rng, majorant);
            currDist += t;
            if (t > d || currDist > maxDist) break;
            vec3 xNew = x + w * t;
            float density = sampleCloud(c, rng, xNew);
            x = xNew;
            float Pn = (majorant - c.sc->extinction[0] * density) / majorant;
            float xi = rng.random();
            if (xi < 1 - Pn) return 0.0f;
            d -= t;
        }
    }
    return 1.0f;
}

float transmittanceResidualRatio(const Ctx& c, Rng& rng, vec3 x, vec3 w, float maxDist) {
    // Supervoxel walk of ResidualRatioTracking.glsl:101-194 with the estimator product T_c * T_r of :33-76,
    // each supervoxel interval clipped to [0, maxDist] from the query origin.
    const vec3 EPSILON_VEC(1e-6f);
    float tMinVal, tMaxVal;
    if (!rayBoxIntersect(c.boxMin + EPSILON_VEC, c.boxMax - EPSILON_VEC, x, w, tMinVal, tMaxVal)) return 1.0f;
    vec3 origin = x;
    x += w * tMinVal;
    vec3 oldX = x;
    float dTotal = tMaxVal - tMinVal;
    float limit = maxDist - tMinVal;
    SvDda q;
    if (!svDdaInit(c, x, w, dTotal, q)) return 1.0f;
    float T = 1.0f;
    float rws = 0.0f, rT = 0.0f, rD = 0.0f;
    (void)origin;
    while (svInside(c, q.idx[0], q.idx[1], q.idx[2])) {
        int si = svIndex(c, q.idx[0], q.idx[1], q.idx[2]);
        float mu_c = c.rrt[2 * si], mu_r_bar = c.rrt[2 * si + 1];
        vec3 mn, mx;
        svBounds(c, q.idx, mn, mx);
        float tMinVoxel = 0.0f, tMaxVoxel = 0.0f;
        rayBoxIntersect(mn, mx, oldX, w, tMinVoxel, tMaxVoxel);
        tMaxVoxel = std::min(tMaxVoxel, limit);
        if (tMaxVoxel > tMinVoxel) {
            vec3 xs = oldX + w * tMinVoxel;
            T *= residualRatioTrackingEstimator(c, rng, xs, w, tMinVoxel, tMaxVoxel, T, rws, rT, rD, mu_c, mu_r_bar);
        }
        svDdaAdvance(q);
    }
    return T;
}

float transmittanceDecomposition(const Ctx& c, Rng& rng, vec3 x, vec3 w, float maxDist) {
    // Free-flight part of DecompositionTracking.glsl:35-123: any accepted control or residual collision ends it.
    const vec3 EPSILON_VEC(1e-6f);
    float tMinVal, tMaxVal;
    if (!rayBoxIntersect(c.boxMin + EPSILON_VEC, c.boxMax - EPSILON_VEC, x, w, tMinVal, tMaxVal)) return 1.0f;
    float majorant = c.sc->extinction[0];
    vec3 voxelGridSize(float(c.dim[0]), float(c.dim[1]), float(c.dim[2]));
    vec3 boxDelta = c.boxMax - c.boxMin;
    vec3 superVoxelSize = vec3(float(c.svSize)) * boxDelta / voxelGridSize;
    x += w * tMinVal;
    float travelled = tMinVal;
    vec3 startPoint = (x - c.boxMin) / boxDelta * voxelGridSize / vec3(float(c.svSize));
    int idx[3] = {int(std::floor(startPoint.x)), int(std::floor(startPoint.y)), int(std::floor(startPoint.z))};
    while (svInside(c, idx[0], idx[1], idx[2])) {
        vec3 minP = c.boxMin + vec3(float(idx[0]), float(idx[1]), float(idx[2])) * superVoxelSize;
        vec3 maxP = minP + superVoxelSize;
        float tMinSV = 0.0f, tMaxSV = 0.0f;
        rayBoxIntersect(minP, maxP, x, w, tMinSV, tMaxSV);
        float d_max = tMaxSV - tMinSV;
        x += w * tMinSV;
        travelled += tMinSV;
        float remaining = maxDist - travelled;
        if (remaining <= 0.0f) break;
        float d_use = std::min(d_max, remaining);
        int si = svIndex(c, idx[0], idx[1], idx[2]);
        float mmMin = c.decomp[2 * si], mmMax = c.decomp[2 * si + 1];
        if (mmMax >= 1e-5f) {
            float mu_c_t = std::max(0.0000000001f, majorant * mmMin);
            float majorant_r_local = std::max(0.0000000001f, majorant * mmMax - mu_c_t);
            float t_c = freePath(rng, mu_c_t);
            float t_r = 0.0f;
            while (true) {
                t_r -= std::log(std::max(0.0000000001f, 1 - rng.random())) / majorant_r_local;
                if (t_c >= d_use && t_r >= d_use) break;
                if (t_c <= t_r) return 0.0f;
                vec3 xs = x + w * t_r;
                float density = sampleCloud(c, rng, xs);
                if (rng.random() * majorant_r_local < c.sc->extinction[0] * density - mu_c_t) return 0.0f;
            }
        }
        if (d_use < d_max) break;
        x = x + d_max * w;
        travelled += d_max;
        vec3 cellCenter = (minP + maxP) * 0.5f;
        vec3 mov = x + w * 0.00001f - cellCenter;
        vec3 smov(signf(mov.x), signf(mov.y), signf(mov.z));
        mov = mov * smov;
        idx[0] += ((mov.x >= mov.y && mov.x >= mov.z) ? 1 : 0) * int(smov.x);
        idx[1] += ((mov.y >= mov.x && mov.y >= mov.z) ? 1 : 0) * int(smov.y);
        idx[2] += ((mov.z >= mov.x && mov.z >= mov.y) ? 1 : 0) * int(smov.z);
    }
    return 1.0f;
}

/* ---------------- FP64 exact optical depth ---------------- */
struct D3 { double v[3]; };

inline double normVoxelD(const Ctx& c, long x, long y, long z) {
    double raw = 0.0;
    if (x >= 0 && y >= 0 && z >= 0 && x < c.dim[0] && y < c.dim[1] && z < c.dim[2]) {
        raw = double(c.sc->density[size_t(x) + (size_t(y) + size_t(z) * c.dim[1]) * c.dim[0]]);
    }
    return (raw - double(c.vmin)) / (double(c.vmax) - double(c.vmin));
}

double opticalDepthD(const Ctx& c, const double o[3], const double d[3], double tmaxQuery) {
    // Clip the segment to the box (slab test in FP64).
    double t0 = 0.0, t1 = tmaxQuery;
    double bmin[3] = {c.boxMin.x, c.boxMin.y, c.boxMin.z}, bmax[3] = {c.boxMax.x, c.boxMax.y, c.boxMax.z};
    for (int a = 0; a < 3; a++) {
        if (std::fabs(d[a]) < 1e-300) {
            if (o[a] < bmin[a] || o[a] > bmax[a]) return 0.0;
            continue;
        }
        double ta = (bmin[a] - o[a]) / d[a], tb = (bmax[a] - o[a]) / d[a];
        if (ta > tb) std::swap(ta, tb);
        t0 = std::max(t0, ta);
        t1 = std::min(t1, tb);
    }
    if (t1 <= t0) return 0.0;
    // Texel-space affine map u(t) = A + B t.
    double A[3], B[3];
    for (int a = 0; a < 3; a++) {
        double scale = double(c.dim[a]) / (bmax[a] - bmin[a]);
        A[a] = (o[a] - bmin[a]) * scale;
        B[a] = d[a] * scale;
    }
    bool nearest = c.sc->interpolation == OR_NEAREST;
    double shift = nearest ? 0.0 : 0.5;  // trilinear lattice at voxel centres
    // Break points where any coordinate crosses an integer lattice plane.
    std::vector<double> ts;
    ts.push_back(t0);
    ts.push_back(t1);
    for (int a = 0; a < 3; a++) {
        if (std::fabs(B[a]) < 1e-300) continue;
        double u0 = A[a] + B[a] * t0 - shift, u1 = A[a] + B[a] * t1 - shift;
        double lo = std::min(u0, u1), hi = std::max(u0, u1);
        for (double k = std::ceil(lo); k <= hi; k += 1.0) {
            double t = (k + shift - A[a]) / B[a];
            if (t > t0 && t < t1) ts.push_back(t);
        }
    }
    std::sort(ts.begin(), ts.end());
    double sigma = double(c.sc->extinction[0]);
    double tau = 0.0;
    const double g = 1.0 / std::sqrt(3.0);
    for (size_t i = 0; i + 1 < ts.size(); i++) {
        double ta = ts[i], tb = ts[i + 1];
        if (tb <= ta) continue;
        double tm = 0.5 * (ta + tb), h = 0.5 * (tb - ta);
        if (nearest) {
            long ix = long(std::floor(A[0] + B[0] * tm));
            long iy = long(std::floor(A[1] + B[1] * tm));
            long iz = long(std::floor(A[2] + B[2] * tm));
            tau += sigma * normVoxelD(c, ix, iy, iz) * (tb - ta);
        } else {
            // Cell lattice index from the midpoint, then evaluate the trilinear interpolant at two GL nodes.
            double um[3];
            long base[3];
            for (int a = 0; a < 3; a++) {
                um[a] = A[a] + B[a] * tm - 0.5;
                base[a] = long(std::floor(um[a]));
            }
            double corner[8];
            for (int k = 0; k < 8; k++) {
                corner[k] = normVoxelD(c, base[0] + (k & 1), base[1] + ((k >> 1) & 1), base[2] + ((k >> 2) & 1));
            }
            double acc = 0.0;
            for (int n = 0; n < 2; n++) {
                double t = tm + (n == 0 ? -g : g) * h;
                double f[3];
                for (int a = 0; a < 3; a++) f[a] = (A[a] + B[a] * t - 0.5) - double(base[a]);
                double v = 0.0;
                for (int k = 0; k < 8; k++) {
                    double wgt = ((k & 1) ? f[0] : 1 - f[0]) * (((k >> 1) & 1) ? f[1] : 1 - f[1]) *
                                 (((k >> 2) & 1) ? f[2] : 1 - f[2]);
                    v += wgt * corner[k];
                }
                acc += v;
            }
            tau += sigma * acc * h;
        }
    }
    return tau;
}

}  // namespace

/* ================================ C ABI ================================ */

extern "C" void oracle_hybrid_taus(uint32_t seed, int n, float* out) {
    Rng r;
    r.init(seed);
    for (int i = 0; i < n; i++) out[i] = r.random();
}

extern "C" void oracle_camera_matrices(const double pos[3], const double target[3], const double up[3], double fovy,
                                       double aspect, double znear, double zfar, float inv_view_proj[16],
                                       float inv_view[16], float view_proj[16]) {
    // View: right-handed look-at. Projection: right-handed perspective with depth in [0,1] and the y axis flipped
    // so that image row 0 is the top of the view (Vulkan-style clip space).
    double f[3] = {target[0] - pos[0], target[1] - pos[1], target[2] - pos[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& v : f) v /= fl;
    double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double& v : s) v /= sl;
    double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    double V[16] = {0};
    V[0] = s[0]; V[4] = s[1]; V[8] = s[2];
    V[1] = u[0]; V[5] = u[1]; V[9] = u[2];
    V[2] = -f[0]; V[6] = -f[1]; V[10] = -f[2];
    V[12] = -(s[0] * pos[0] + s[1] * pos[1] + s[2] * pos[2]);
    V[13] = -(u[0] * pos[0] + u[1] * pos[1] + u[2] * pos[2]);
    V[14] = (f[0] * pos[0] + f[1] * pos[1] + f[2] * pos[2]);
    V[15] = 1.0;
    double th = std::tan(fovy / 2.0);
    double P[16] = {0};
    P[0] = 1.0 / (aspect * th);
    P[5] = -1.0 / th;
    P[10] = zfar / (znear - zfar);
    P[11] = -1.0;
    P[14] = -(zfar * znear) / (zfar - znear);
    double M[16];
    for (int col = 0; col < 4; col++)
        for (int row = 0; row < 4; row++) {
            double acc = 0;
            for (int k = 0; k < 4; k++) acc += P[k * 4 + row] * V[col * 4 + k];
            M[col * 4 + row] = acc;
        }
    // Inverse by Gauss-Jordan in FP64.
    auto invert = [](const double* in, double* out) {
        double a[4][8];
        for (int r = 0; r < 4; r++)
            for (int cc = 0; cc < 4; cc++) {
                a[r][cc] = in[cc * 4 + r];
                a[r][cc + 4] = (r == cc) ? 1.0 : 0.0;
            }
        for (int col = 0; col < 4; col++) {
            int piv = col;
            for (int r = col + 1; r < 4; r++)
                if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
            for (int cc = 0; cc < 8; cc++) std::swap(a[col][cc], a[piv][cc]);
            double p = a[col][col];
            for (int cc = 0; cc < 8; cc++) a[col][cc] /= p;
            for (int r = 0; r < 4; r++) {
                if (r == col) continue;
                double fct = a[r][col];
                for (int cc = 0; cc < 8; cc++) a[r][cc] -= fct * a[col][cc];
            }
        }
        for (int r = 0; r < 4; r++)
            for (int cc = 0; cc < 4; cc++) out[cc * 4 + r] = a[r][cc + 4];
    };
    double IM[16], IV[16];
    invert(M, IM);
    invert(V, IV);
    for (int i = 0; i < 16; i++) {
        inv_view_proj[i] = float(IM[i]);
        inv_view[i] = float(IV[i]);
        if (view_proj) view_proj[i] = float(M[i]);
    }
}

namespace {
/* One sample of pixel (x, y) in frame f: the seed and camera jitter of Clouds.glsl:73-90, then the estimator. */
vec3 renderSample(const Ctx& c, const OracleScene* scene, int width, int height, int x, int y, int f) {
    uint32_t frameGlobal = uint32_t(f);
    uint32_t seed = frameGlobal * uint32_t(width) * uint32_t(height) + uint32_t(x) + uint32_t(y) * uint32_t(width);
    seed += scene->custom_seed_offset;
    Rng rng;
    rng.init(seed);
    float r0 = rng.random();
    float r1 = rng.random();
    float sx = 2.0f * (float(x) + r0) / float(width) - 1;
    float sy = 2.0f * (float(y) + r1) / float(height) - 1;
    vec3 xo, w;
    createCameraRay(c, sx, sy, xo, w);
    return tracePath(c, rng, xo, w);
}
}  // namespace

extern "C" int oracle_render(const OracleScene* scene, int width, int height, int frame_start, int num_frames,
                             float* acc_rgb, double* sum_sq, int num_threads) {
    if (!scene || !acc_rgb || width <= 0 || height <= 0 || num_frames < 0) return 1;
    Ctx c;
    initCtx(c, scene);
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int y = 0; y < height; y++) {
        for (int x = 0; x < width; x++) {
            float* acc = acc_rgb + (size_t(y) * width + x) * 3;
            for (int f = frame_start; f < frame_start + num_frames; f++) {
                const uint32_t frame = uint32_t(f);
                vec3 result = renderSample(c, scene, width, height, x, y, f);
                if (sum_sq) {
                    double* q = sum_sq + (size_t(y) * width + x) * 3;
                    q[0] += double(result.x) * result.x;
                    q[1] += double(result.y) * result.y;
                    q[2] += double(result.z) * result.z;
                }
                vec3 resultOld = frame == 0 ? vec3(0.0f) : vec3(acc[0], acc[1], acc[2]);
                result = mix3(resultOld, result, 1.0f / float(frame + 1));
                acc[0] = result.x;
                acc[1] = result.y;
                acc[2] = result.z;
            }
        }
    }
    return 0;
}

extern "C" int oracle_render_sums(const OracleScene* scene, int width, int height, int frame_start, int num_frames,
                                  double* sum_rgb, double* sum_sq, int num_threads) {
    if (!scene || !sum_rgb || !sum_sq || width <= 0 || height <= 0 || num_frames < 0) return 1;
    Ctx c;
    initCtx(c, scene);
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int y = 0; y < height; y++) {
        for (int x = 0; x < width; x++) {
            double* m = sum_rgb + (size_t(y) * width + x) * 3;
            double* q = sum_sq + (size_t(y) * width + x) * 3;
            for (int f = frame_start; f < frame_start + num_frames; f++) {
                const vec3 r = renderSample(c, scene, width, height, x, y, f);
                m[0] += double(r.x);
                m[1] += double(r.y);
                m[2] += double(r.z);
                q[0] += double(r.x) * r.x;
                q[1] += double(r.y) * r.y;
                q[2] += double(r.z) * r.z;
            }
        }
    }
    return 0;
}

namespace {
/* getCloudFiniteDifference, dense branch (VptUtils.glsl:726-748): central differences of the normalised grid one
 * texel apart around the (stochastically jittered) texture coordinate, scaled by 100 / dim and normalised. */
vec3 cloudFiniteDifference(const Ctx& c, Rng& rng, vec3 pos) {
    vec3 coord = (pos - c.boxMin) / (c.boxMax - c.boxMin);
    const vec3 dim(float(c.dim[0]), float(c.dim[1]), float(c.dim[2]));
    if (c.sc->interpolation == OR_STOCHASTIC) {
        float r0 = rng.random();
        float r1 = rng.random();
        float r2 = rng.random();
        coord += vec3(r0 - 0.5f, r1 - 0.5f, r2 - 0.5f) / dim;
    }
    auto S = [&](vec3 q) { return (textureFetch(c, q) - c.vmin) / (c.vmax - c.vmin); };
    vec3 d = vec3(S(coord - vec3(1, 0, 0) / dim) - S(coord + vec3(1, 0, 0) / dim),
                  S(coord - vec3(0, 1, 0) / dim) - S(coord + vec3(0, 1, 0) / dim),
                  S(coord - vec3(0, 0, 1) / dim) - S(coord + vec3(0, 0, 1) / dim)) / dim * 100.0f;
    const float l = length(d);
    if (l > 1e-5f) d = d / l;
    return d;
}
inline void store(float* img, size_t px, int ch, const float* v) {
    if (img)
        for (int i = 0; i < ch; i++) img[px * ch + i] = v[i];
}
}  // namespace

extern "C" int oracle_render_features(const OracleScene* scene, int width, int height, int frame_start,
                                      int num_frames, const float prev_view_proj[16], float far_distance,
                                      OracleFeatureMaps* m, int num_threads) {
    if (!scene || !m || width <= 0 || height <= 0 || num_frames < 0) return 1;
    Ctx c;
    initCtx(c, scene);
    const vec3 albedo(scene->scattering_albedo);
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int y = 0; y < height; y++) {
        for (int x = 0; x < width; x++) {
            const size_t px = size_t(y) * width + x;
            for (int f = frame_start; f < frame_start + num_frames; f++) {
                // Clouds.glsl:73-383 for one sample with the enabled feature maps.
                const uint32_t frame = uint32_t(f);
                uint32_t seed = frame * uint32_t(width) * uint32_t(height) + uint32_t(x) + uint32_t(y) * uint32_t(width);
                seed += scene->custom_seed_offset;
                Rng rng;
                rng.init(seed);
                float r0 = rng.random();
                float r1 = rng.random();
                float sx = 2.0f * (float(x) + r0) / float(width) - 1;
                float sy = 2.0f * (float(y) + r1) / float(height) - 1;
                vec3 xo, w;
                createCameraRay(c, sx, sy, xo, w);
                ScatterEvent fe;
                vec3 result = tracePath(c, rng, xo, w, &fe);
                const float a = 1.0f / float(frame + 1);
                auto mix4 = [&](float* img, const float* v) {
                    float o[4] = {0, 0, 0, 0};
                    if (frame != 0)
                        for (int i = 0; i < 4; i++) o[i] = img[px * 4 + i];
                    for (int i = 0; i < 4; i++) o[i] = mixf(o[i], v[i], a);
                    store(img, px, 4, o);
                };
                if (m->result) {
                    const float v[4] = {result.x, result.y, result.z, 1.0f};
                    mix4(m->result, v);
                }
                if (m->cloud_only) {
                    const float v[4] = {fe.hasValue ? result.x : 0.0f, fe.hasValue ? result.y : 0.0f,
                                        fe.hasValue ? result.z : 0.0f, fe.hasValue ? 1.0f : 0.0f};
                    mix4(m->cloud_only, v);
                }
                if (m->background) {
                    const vec3 b = fe.hasValue ? sampleSkybox(c, w) : result;
                    const float v[4] = {b.x, b.y, b.z, 1.0f};
                    mix4(m->background, v);
                }
                if (m->first_x) {
                    const float v[4] = {fe.hasValue ? fe.x.x : 0.0f, fe.hasValue ? fe.x.y : 0.0f,
                                        fe.hasValue ? fe.x.z : 0.0f, fe.hasValue ? 1.0f : 0.0f};
                    mix4(m->first_x, v);
                }
                if (m->depth) {  // (mean, std) with the second moment rebuilt from the stored std
                    float o[2] = {0, 0};
                    if (frame != 0) {
                        o[0] = m->depth[px * 2];
                        o[1] = m->depth[px * 2 + 1] * m->depth[px * 2 + 1] + o[0] * o[0];
                    }
                    const float v0 = fe.hasValue ? fe.depth : 0.0f, v1 = fe.hasValue ? fe.depth * fe.depth : 0.0f;
                    const float n0 = mixf(o[0], v0, a), n1 = mixf(o[1], v1, a);
                    const float v[2] = {n0, std::sqrt(std::max(0.0f, n1 - n0 * n0))};
                    store(m->depth, px, 2, v);
                }
                if (m->depth_far) {
                    const float o = frame == 0 ? 0.0f : m->depth_far[px];
                    const float v = mixf(o, fe.hasValue ? fe.depth : far_distance, a);
                    store(m->depth_far, px, 1, &v);
                }
                if (m->density) {
                    float o[2] = {0, 0};
                    if (frame != 0) {
                        o[0] = m->density[px * 2];
                        o[1] = m->density[px * 2 + 1] * m->density[px * 2 + 1] + o[0] * o[0];
                    }
                    const float v0 = fe.hasValue ? fe.density * 0.001f : 0.0f;
                    const float v1 = fe.hasValue ? fe.density * fe.density * 0.001f * 0.001f : 0.0f;
                    const float n0 = mixf(o[0], v0, a), n1 = mixf(o[1], v1, a);
                    const float v[2] = {n0, std::sqrt(std::max(0.0f, n1 - n0 * n0))};
                    store(m->density, px, 2, v);
                }
                if (m->reproj_uv) {
                    float v[2] = {-1.0f, -1.0f};
                    if (frame != 0) {
                        v[0] = m->reproj_uv[px * 2];
                        v[1] = m->reproj_uv[px * 2 + 1];
                    }
                    if (fe.hasValue) {
                        const vec4 pc = mul(prev_view_proj, vec4{fe.x.x, fe.x.y, fe.x.z, 1.0f});
                        v[0] = pc.x / pc.w * 0.5f + 0.5f;
                        v[1] = pc.y / pc.w * 0.5f + 0.5f;
                    }
                    store(m->reproj_uv, px, 2, v);
                }
                if (m->albedo) {
                    const float v[4] = {fe.hasValue ? albedo.x : 0.0f, fe.hasValue ? albedo.y : 0.0f,
                                        fe.hasValue ? albedo.z : 0.0f, fe.hasValue ? 1.0f : 0.0f};
                    mix4(m->albedo, v);
                }
                float nabla[2] = {0.0f, 0.0f};
                const bool wantNormal = m->normal || m->depth_nabla || m->depth_fwidth;
                if (fe.hasValue) {
                    if (wantNormal) {
                        vec3 diff = cloudFiniteDifference(c, rng, fe.x);
                        vec3 old(0.0f);
                        if (frame != 0 && m->normal) old = vec3(m->normal[px * 4], m->normal[px * 4 + 1], m->normal[px * 4 + 2]);
                        diff = mix3(old, diff, 1.0f / (float(frame) + 1.0f));
                        const float v[4] = {diff.x, diff.y, diff.z, 1.0f};
                        store(m->normal, px, 4, v);
                        // inverseTransposedViewMatrix * (diff, 0) = transpose(inverseView) * (diff, 0)
                        const float* iv = scene->inverse_view;
                        const float A = iv[0] * diff.x + iv[1] * diff.y + iv[2] * diff.z;
                        const float B = iv[4] * diff.x + iv[5] * diff.y + iv[6] * diff.z;
                        nabla[0] = A / std::sqrt(1.0f - A * A);
                        nabla[1] = B / std::sqrt(1.0f - B * B);
                    }
                    const float v[4] = {fe.w.x, fe.w.y, fe.w.z, fe.pdf_w};
                    store(m->first_w, px, 4, v);
                } else {
                    if (m->normal) {
                        const float k = float(frame) / (float(frame) + 1.0f);
                        const float v[4] = {frame == 0 ? 0.0f : m->normal[px * 4] * k,
                                            frame == 0 ? 0.0f : m->normal[px * 4 + 1] * k,
                                            frame == 0 ? 0.0f : m->normal[px * 4 + 2] * k, 1.0f};
                        store(m->normal, px, 4, v);
                    }
                    const float z4[4] = {0, 0, 0, 0};
                    store(m->first_w, px, 4, z4);
                }
                if (m->flow) {
                    float v[2] = {0.0f, 0.0f};
                    if (fe.hasValue) {
                        vec4 q = mul(prev_view_proj, vec4{fe.x.x, fe.x.y, fe.x.z, 1.0f});
                        const float lx = (0.5f * (q.x / q.w) + 0.5f) * float(width) - 0.5f;
                        const float ly = (0.5f * (q.y / q.w) + 0.5f) * float(height) - 0.5f;
                        v[0] = float(x) - lx;
                        v[1] = float(y) - ly;
                    }
                    store(m->flow, px, 2, v);
                }
                if (m->flow_reverse) {
                    const vec4 wp = fe.hasValue ? vec4{fe.x.x, fe.x.y, fe.x.z, 1.0f} : vec4{w.x, w.y, w.z, 0.0f};
                    vec4 q = mul(prev_view_proj, wp);
                    const float lx = (0.5f * (q.x / q.w) + 0.5f) * float(width) - 0.5f;
                    const float ly = (0.5f * (q.y / q.w) + 0.5f) * float(height) - 0.5f;
                    const float v[2] = {lx - float(x), ly - float(y)};
                    store(m->flow_reverse, px, 2, v);
                }
                if (m->depth_nabla) {
                    if (frame != 0) {
                        nabla[0] = mixf(m->depth_nabla[px * 2], nabla[0], a);
                        nabla[1] = mixf(m->depth_nabla[px * 2 + 1], nabla[1], a);
                    }
                    store(m->depth_nabla, px, 2, nabla);
                }
                if (m->depth_fwidth) {
                    float v = std::fabs(nabla[0]) + std::fabs(nabla[1]);
                    if (frame != 0) v = mixf(m->depth_fwidth[px], v, a);
                    store(m->depth_fwidth, px, 1, &v);
                }
            }
        }
    }
    return 0;
}

extern "C" int oracle_transmittance(const OracleScene* scene, int estimator, int n_rays, const float* o,
                                    const float* d, const float* t_max, int n_samples, uint32_t seed, double* mean,
                                    double* m2, int num_threads) {
    if (!scene || n_rays < 0 || n_samples <= 0) return 1;
    Ctx c;
    initCtx(c, scene);
    if (estimator == OR_T_RESIDUAL_RATIO || estimator == OR_T_DECOMPOSITION) {
        if (c.rrt.empty()) buildSupervoxels(c, scene->extinction[0]);
    }
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int i = 0; i < n_rays; i++) {
        vec3 x(o + 3 * size_t(i)), w(d + 3 * size_t(i));
        float lim = t_max ? t_max[i] : 3.0e38f;
        double s1 = 0.0, s2 = 0.0;
        for (int s = 0; s < n_samples; s++) {
            Rng rng;
            rng.init(seed + uint32_t(i) * uint32_t(n_samples) + uint32_t(s));
            float T;
            switch (estimator) {
                case OR_T_RATIO: T = calculateTransmittanceDistance(c, rng, x, w, lim, true); break;
                case OR_T_RESIDUAL_RATIO: T = transmittanceResidualRatio(c, rng, x, w, lim); break;
                case OR_T_DECOMPOSITION: T = transmittanceDecomposition(c, rng, x, w, lim); break;
                default: T = transmittanceDelta(c, rng, x, w, lim); break;
            }
            s1 += T;
            s2 += double(T) * T;
        }
        mean[i] = s1 / n_samples;
        if (m2) m2[i] = s2 / n_samples;
    }
    return 0;
}

extern "C" double oracle_optical_depth(const OracleScene* scene, const double o[3], const double d[3], double t_max) {
    Ctx c;
    c.sc = scene;
    c.boxMin = vec3(scene->box_min);
    c.boxMax = vec3(scene->box_max);
    c.dim[0] = scene->nx; c.dim[1] = scene->ny; c.dim[2] = scene->nz;
    c.vmin = scene->voxel_value_min;
    c.vmax = scene->voxel_value_max;
    if (c.vmin == c.vmax) c.vmin = 0.0f;
    return opticalDepthD(c, o, d, t_max);
}

extern "C" int oracle_optical_depth_batch(const OracleScene* scene, int n_rays, const float* o, const float* d,
                                          const float* t_max, double* tau, int num_threads) {
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int i = 0; i < n_rays; i++) {
        double oo[3] = {o[3 * i], o[3 * i + 1], o[3 * i + 2]};
        double dd[3] = {d[3 * i], d[3 * i + 1], d[3 * i + 2]};
        tau[i] = oracle_optical_depth(scene, oo, dd, t_max ? double(t_max[i]) : 1e30);
    }
    return 0;
}

extern "C" int oracle_supervoxel_tables(const OracleScene* scene, float extinction_x, int* grid_size3,
                                        float* rrt_pairs, float* decomp_pairs, uint8_t* occupancy) {
    Ctx c;
    c.sc = scene;
    c.boxMin = vec3(scene->box_min);
    c.boxMax = vec3(scene->box_max);
    c.dim[0] = scene->nx; c.dim[1] = scene->ny; c.dim[2] = scene->nz;
    c.vmin = scene->voxel_value_min;
    c.vmax = scene->voxel_value_max;
    if (c.vmin == c.vmax) c.vmin = 0.0f;
    buildSupervoxels(c, extinction_x);
    for (int a = 0; a < 3; a++) grid_size3[a] = c.svGrid[a];
    size_t n = size_t(c.svGrid[0]) * c.svGrid[1] * c.svGrid[2];
    if (rrt_pairs) std::memcpy(rrt_pairs, c.rrt.data(), 2 * n * sizeof(float));
    if (decomp_pairs) std::memcpy(decomp_pairs, c.decomp.data(), 2 * n * sizeof(float));
    if (occupancy) std::memcpy(occupancy, c.occ.data(), n);
    return c.svSize;
}

extern "C" int oracle_trace_rays(const OracleScene* scene, int n_rays, const float* o, const float* d, uint32_t seed,
                                 int rrt_max_iterations, float* out_rgb, int* out_iterations, int num_threads) {
    if (!scene || n_rays < 0 || !o || !d || !out_rgb) return 1;
    Ctx base;
    initCtx(base, scene);
    if (rrt_max_iterations >= 0) base.rrtMaxIterations = rrt_max_iterations;
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int i = 0; i < n_rays; i++) {
        t_lastRrtIterations = 0;
        Rng rng;
        rng.init(seed + uint32_t(i));
        const vec3 r = tracePath(base, rng, vec3(o + 3 * size_t(i)), vec3(d + 3 * size_t(i)));
        const int iterations = t_lastRrtIterations;
        out_rgb[3 * size_t(i)] = r.x;
        out_rgb[3 * size_t(i) + 1] = r.y;
        out_rgb[3 * size_t(i) + 2] = r.z;
        if (out_iterations) out_iterations[i] = iterations;
    }
    return 0;
}

extern "C" int oracle_sky(const OracleScene* scene, int n, const float* dirs, float* out_rgb) {
    if (!scene || n < 0 || !dirs || !out_rgb) return 1;
    Ctx c;
    c.sc = scene;
    for (int i = 0; i < n; i++) {
        const vec3 w(dirs + 3 * size_t(i));
        const vec3 r = sampleSkybox(c, w) + sampleLight(c, w);
        out_rgb[3 * size_t(i)] = r.x;
        out_rgb[3 * size_t(i) + 1] = r.y;
        out_rgb[3 * size_t(i) + 2] = r.z;
    }
    return 0;
}

extern "C" int oracle_light_radiance(const OracleScene* scene, int light_idx, int n, const float* positions,
                                     float* out_rgb) {
    if (!scene || light_idx < 0 || light_idx >= scene->num_lights || n < 0 || !positions || !out_rgb) return 1;
    Ctx c;
    c.sc = scene;
    for (int i = 0; i < n; i++) {
        const vec3 r = sampleHeadlight(c, vec3(positions + 3 * size_t(i)), uint32_t(light_idx));
        out_rgb[3 * size_t(i)] = r.x;
        out_rgb[3 * size_t(i) + 1] = r.y;
        out_rgb[3 * size_t(i) + 2] = r.z;
    }
    return 0;
}
