// Device-side path tracer shared by the render kernels: the cell walker, the per-lane path state machine of every
// estimator, and the helpers the persistent path-pool kernels (single render and batched queue) are built from.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include "device_common.hpp"
#include "nanovdb_access.hpp"
#include "render_api.hpp"

namespace crvpt {

constexpr float kInf = 3.0e38f;


// ------------------------------------------------------------------------------------------------------------
// Cell walker
//
// A ray segment is walked cell by cell through a grid of supervoxels with a piecewise-constant collision rate:
//  * delta, ratio and next-event tracking (and the transmittance queries) walk the majorant grid: rate
//    extinction * (max density reachable by a lookup in the cell);
//  * decomposition tracking walks the reference's (min, max) table: control rate mu_c = extinction * min and
//    residual rate mu_r = extinction * max - mu_c (DecompositionTracking.glsl:67-75). The competing control and
//    residual exponentials of the reference are one Poisson process of rate mu_c + mu_r whose events are control
//    collisions with probability mu_c / (mu_c + mu_r);
//  * residual ratio tracking walks the reference's (mu_c, mu_r_bar) table: tentative collisions at rate mu_r_bar,
//    and the control transmittance exp(-integral of mu_c) accumulated along the way (ResidualRatioTracking.glsl:
//    33-76).
// An Exp(1) optical budget carries over cell boundaries (the process is memoryless, so this samples the same
// collision distribution as the reference's fresh draw per supervoxel). One call crosses up to maxCrossings<MODE>()
// cells before it returns, so empty space costs cell reads but no random-number blocks.
// ------------------------------------------------------------------------------------------------------------

// Per-estimator launch tuning, measured on MI355X (1024^2 renders of the 512^3 sparse grid, 64-256 spp, and the
// dense 128^3 next-event workload): cell crossings per walker call (more crossings per step means fewer steps but
// more lanes idle while the others cross), and the waves per SIMD the pool kernels are compiled for (128 registers
// per lane at 4 waves, 96 at 5, 80 at 6). Next-event tracking carries the most per-lane state (the shadow ray's):
// spectral next-event tracking runs best at 4 waves, next-event and spectral delta tracking at 5 (dense 128^3
// next-event workload: 51.5 ms against 54.7 ms at 4 waves), residual ratio and decomposition tracking at 6 (residual
// ratio: 351 ms against 388 ms at 5 waves and 410 ms at 7 per 2^28-path launch), delta and ratio tracking at 7 (66.4
// and 124.0 ms against 67.4 and 127.5 ms at 6); residual ratio and next-event tracking at 2 crossings, the others at
// 3. The batched queue's kernel also holds its request's constants per lane and keeps next-event tracking at 4 waves,
// residual ratio and decomposition tracking at 5, delta and ratio tracking at 6 (batchPoolWaves).
template <int MODE>
constexpr int maxCrossings() {
    return (MODE == VPT_RESIDUAL_RATIO_TRACKING || MODE == VPT_NEXT_EVENT_TRACKING ||
            MODE == VPT_NEXT_EVENT_TRACKING_SPECTRAL) ? 2 : 3;
}
template <int MODE>
constexpr int poolWaves() {
    if (MODE == VPT_NEXT_EVENT_TRACKING_SPECTRAL) return 4;
    if (MODE == VPT_NEXT_EVENT_TRACKING) return 5;
    if (MODE == VPT_DELTA_TRACKING || MODE == VPT_RATIO_TRACKING) return 7;
    return MODE == VPT_SPECTRAL_DELTA_TRACKING ? 5 : 6;
}

template <int MODE>
constexpr int batchPoolWaves() {
    if (MODE == VPT_NEXT_EVENT_TRACKING) return 4;
    if (MODE == VPT_RESIDUAL_RATIO_TRACKING || MODE == VPT_DECOMPOSITION_TRACKING) return 5;
    if (MODE == VPT_DELTA_TRACKING || MODE == VPT_RATIO_TRACKING) return 6;
    return poolWaves<MODE>();
}

template <int MODE>
constexpr bool walksReferenceTable() {
    return MODE == VPT_DECOMPOSITION_TRACKING || MODE == VPT_RESIDUAL_RATIO_TRACKING;
}
template <int MODE>
constexpr bool isSpectral() {
    return MODE == VPT_SPECTRAL_DELTA_TRACKING || MODE == VPT_NEXT_EVENT_TRACKING_SPECTRAL;
}
template <int MODE>
constexpr bool isNee() {
    return MODE == VPT_NEXT_EVENT_TRACKING || MODE == VPT_NEXT_EVENT_TRACKING_SPECTRAL;
}

struct Walker {
    uint32_t svReads;   // cell reads since the launch started (work counter)
    V3 o, w;            // ray origin / direction (world)
    float t, tEnd;      // current distance along the ray, end of the segment
    int ix, iy, iz;     // current cell
    float nx, ny, nz;   // distance of the next cell plane per axis (the spacing, cell edge / |w_a|, is recomputed
                        // at each crossing rather than held: the path kernels run at the register cap)
    float mu;           // tentative-collision rate of the current cell, extinction units
    float muC;          // control rate of the current cell (decomposition and residual ratio tracking)
    float ctrl;         // control optical depth walked since the caller last cleared it (residual ratio tracking)
    float ext;          // extinction scale of the segment's majorants (spectral trackers: max component on the
                        // main path, extinction.x on shadow rays)
    float tau;          // remaining optical length of the current Exp(1) sample; < 0 when none is drawn
};

CRVPT_DEV float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

template <int MODE>
CRVPT_DEV float cellWorld(const SupervoxelGrid& S, int a) {
    return walksReferenceTable<MODE>() ? S.refCellWorld[a] : S.cellWorld[a];
}
template <int MODE>
CRVPT_DEV float cellsPerWorld(const SupervoxelGrid& S, int a) {
    return walksReferenceTable<MODE>() ? S.refCellsPerWorld[a] : S.cellsPerWorld[a];
}
template <int MODE>
CRVPT_DEV int cellGridDim(const SupervoxelGrid& S, int a) {
    return walksReferenceTable<MODE>() ? S.refGrid[a] : S.grid[a];
}

// Rates of the cell (W.ix, W.iy, W.iz). Cell indices are 32-bit (the host keeps both cell grids below 2^32 cells):
// the 64-bit index arithmetic cost 2-5 % on the 512^3 sparse sweep.
template <int MODE>
CRVPT_DEV void cellRates(Walker& W, const SceneParams& P, const SupervoxelGrid& S) {
    if (walksReferenceTable<MODE>()) {
        const uint32_t c = (uint32_t(W.iz) * uint32_t(S.refGrid[1]) + uint32_t(W.iy)) * uint32_t(S.refGrid[0]) +
                           uint32_t(W.ix);
        if (MODE == VPT_DECOMPOSITION_TRACKING) {
            const float2 m = reinterpret_cast<const float2*>(S.decomp)[c];
            // supervoxels whose max is below 1e-5 are skipped (DecompositionTracking.glsl:67-69)
            const float muC = fmaxf(1e-10f, P.extinction * m.x);
            const float muR = fmaxf(1e-10f, P.extinction * m.y - muC);
            W.muC = muC;
            W.mu = m.y < 1e-5f ? 0.0f : muC + muR;
        } else {
            const float2 m = reinterpret_cast<const float2*>(S.rrt)[c];
            W.muC = m.x;
            W.mu = m.y;
        }
    } else {
        const float ext = isSpectral<MODE>() ? W.ext : P.extinction;
        W.mu = ext * S.majorant[(uint32_t(W.iz) * uint32_t(S.grid[1]) + uint32_t(W.iy)) * uint32_t(S.grid[0]) +
                                uint32_t(W.ix)];
    }
    W.svReads++;
}

// Starts a segment from x along w, clipped to the box and to [0, limit]. False when the segment is empty.
// The slab test follows rayBoxIntersect (near-zero direction components replaced by 1e-6, tMin >= 0).
//
// Decomposition and residual ratio tracking walk the reference's supervoxel grid until the ray leaves the grid
// (DecompositionTracking.glsl:55-117, ResidualRatioTracking.glsl:111-194), and that grid overhangs the box when the
// grid size is not a multiple of the supervoxel edge: their segments end at the exit from the grid's box, where
// lookups read the zero border (or, with stochastic filtering, the last voxels half a voxel beyond the box). A
// decomposition path that scatters keeps walking the grid (gridBox), residual ratio tracking and camera rays enter
// through the box.
template <int MODE>
CRVPT_DEV bool walkerStart(Walker& W, const SceneParams& P, const SupervoxelGrid& S, V3 x, V3 w, float limit,
                           bool shadow = false, bool gridBox = false) {
    const float ax = fabsf(w.x) <= 0.000001f ? 0.000001f : w.x;
    const float ay = fabsf(w.y) <= 0.000001f ? 0.000001f : w.y;
    const float az = fabsf(w.z) <= 0.000001f ? 0.000001f : w.z;
    const V3 inv = v3(rcpf(ax), rcpf(ay), rcpf(az));
    const V3 hi = walksReferenceTable<MODE>() ? v3(S.refHi[0], S.refHi[1], S.refHi[2]) : v3(P.boxMax);
    const V3 c0 = (v3(P.boxMin) - x) * inv;
    const V3 c1 = ((gridBox ? hi : v3(P.boxMax)) - x) * inv;
    float tMin = fmaxf(fmaxf(fminf(c0.x, c1.x), fminf(c0.y, c1.y)), fminf(c0.z, c1.z));
    tMin = fmaxf(0.0f, tMin);
    float tMax = fminf(fminf(fmaxf(c0.x, c1.x), fmaxf(c0.y, c1.y)), fmaxf(c0.z, c1.z));
    if (tMax <= tMin || tMax <= 0.0f) return false;
    if (walksReferenceTable<MODE>()) {
        const V3 c2 = (hi - x) * inv;
        tMax = fmaxf(tMax, fminf(fminf(fmaxf(c0.x, c2.x), fmaxf(c0.y, c2.y)), fmaxf(c0.z, c2.z)));
    }
    tMax = fminf(tMax, limit);
    if (!(tMax > tMin)) return false;
    W.o = x;
    W.w = w;
    W.t = tMin;
    W.tEnd = tMax;
    W.tau = -1.0f;
    W.ctrl = 0.0f;
    W.ext = (isSpectral<MODE>() && !shadow) ? P.extinctionMax : P.extinction;
    if (!walksReferenceTable<MODE>() && !P.useMajorantGrid) {
        W.mu = isSpectral<MODE>() ? W.ext : P.extinction;
        W.nx = W.ny = W.nz = kInf;
        return true;
    }
    // Cell of a point slightly inside the segment (robust against an entry point on a cell plane).
    const float tc = tMin + 1e-4f * (tMax - tMin);
    const V3 p = x + w * tc;
    int idx[3];
    float nxt[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float wa = comp(w, a), pa = comp(p, a), ia = comp(inv, a);
        const float cw = cellWorld<MODE>(S, a);
        int i = int(floorf((pa - P.boxMin[a]) * cellsPerWorld<MODE>(S, a)));
        i = min(max(i, 0), cellGridDim<MODE>(S, a) - 1);
        idx[a] = i;
        // Components the slab test clamps to 1e-6 never reach a cell plane within the box (a segment moves less
        // than 1e-5 world units along them); the clamped inverse would also carry the wrong sign for tiny negative
        // components.
        if (fabsf(wa) <= 0.000001f) {
            nxt[a] = kInf;
        } else if (wa > 0.0f) {
            nxt[a] = tc + (P.boxMin[a] + float(i + 1) * cw - pa) * ia;
        } else if (wa < 0.0f) {
            nxt[a] = tc + (P.boxMin[a] + float(i) * cw - pa) * ia;
        } else {
            nxt[a] = kInf;
        }
    }
    W.ix = idx[0]; W.iy = idx[1]; W.iz = idx[2];
    W.nx = nxt[0]; W.ny = nxt[1]; W.nz = nxt[2];
    cellRates<MODE>(W, P, S);
    return true;
}

enum : int { WALK_MOVED = 0, WALK_COLLIDE = 1, WALK_END = 2 };

// Distance along the ray between consecutive cell planes of axis a: the cell edge times |1 / w_a| (the same
// operations as walkerStart's plane distances, so the walk is unchanged). Only called for an axis whose next plane
// is finite, i.e. |w_a| > 1e-6.
template <int MODE>
CRVPT_DEV float planeSpacing(const SceneParams& P, const SupervoxelGrid& S, float wa, int a) {
    const float cw = cellWorld<MODE>(S, a);
    const float ia = rcpf(wa);
    return wa > 0.0f ? cw * ia : -cw * ia;
}

// Advances to the next tentative collision (W.t at the collision), the segment end, or, after maxCrossings cell
// crossings without either, returns WALK_MOVED. e is an Exp(1) sample, consumed only when no budget is pending.
template <int MODE>
CRVPT_DEV int walkAdvance(Walker& W, const SceneParams& P, const SupervoxelGrid& S, float e) {
    if (W.tau < 0.0f) W.tau = e;
    for (int crossings = 0;;) {
        const float tExit = fminf(fminf(W.nx, W.ny), fminf(W.nz, W.tEnd));
        const float opt = W.mu * (tExit - W.t);
        if (W.tau < opt) {
            const float dt = W.tau * rcpf(W.mu);
            if (MODE == VPT_RESIDUAL_RATIO_TRACKING) W.ctrl += W.muC * dt;
            W.t += dt;
            W.tau = -1.0f;
            return WALK_COLLIDE;
        }
        W.tau -= opt;
        if (MODE == VPT_RESIDUAL_RATIO_TRACKING) W.ctrl += W.muC * (tExit - W.t);
        W.t = tExit;
        if (tExit >= W.tEnd) return WALK_END;
        if (W.nx <= W.ny && W.nx <= W.nz) {
            W.ix += W.w.x > 0.0f ? 1 : -1;
            W.nx += planeSpacing<MODE>(P, S, W.w.x, 0);
        } else if (W.ny <= W.nz) {
            W.iy += W.w.y > 0.0f ? 1 : -1;
            W.ny += planeSpacing<MODE>(P, S, W.w.y, 1);
        } else {
            W.iz += W.w.z > 0.0f ? 1 : -1;
            W.nz += planeSpacing<MODE>(P, S, W.w.z, 2);
        }
        if ((unsigned)W.ix >= (unsigned)cellGridDim<MODE>(S, 0) ||
            (unsigned)W.iy >= (unsigned)cellGridDim<MODE>(S, 1) ||
            (unsigned)W.iz >= (unsigned)cellGridDim<MODE>(S, 2)) {
            W.tEnd = W.t;
            return WALK_END;
        }
        cellRates<MODE>(W, P, S);
        if (++crossings >= maxCrossings<MODE>()) return WALK_MOVED;
    }
}

// Walker event with the exponential drawn from a buffered stream (transmittance queries).
CRVPT_DEV int walkStep(Walker& W, Rng& rng, const SceneParams& P, const SupervoxelGrid& S) {
    return walkAdvance<VPT_DELTA_TRACKING>(W, P, S, W.tau < 0.0f ? rng.exponential() : 0.0f);
}

template <class Grid>
CRVPT_DEV float density(const Grid& g, const SceneParams& P, Rng& rng, V3 p) {
    return g.lookupWorld(P, rng, p);
}
template <class Grid>
CRVPT_DEV float density(const Grid& g, const SceneParams& P, Rng& rng, V3 p, typename Grid::Cache& gc) {
    return g.lookupWorld(P, rng, p, gc);
}

// ------------------------------------------------------------------------------------------------------------
// Path pool: every estimator as a per-lane state machine advanced one walker event per step.
//   delta tracking          DeltaTracking.glsl:167-345
//   analog ratio tracking   RatioTracking.glsl:25-96
//   next-event tracking     NextEventTracking.glsl:223-394, shadow rays by ratio tracking (NeeTransmittance.glsl)
//   decomposition tracking  DecompositionTracking.glsl:35-123
//   residual ratio tracking ResidualRatioTracking.glsl:33-224 (reservoir-sampled in-scattering, <= 10 iterations)
//
// Random numbers: every step of every lane evaluates one Philox block (counter: step index, pixel, global frame,
// seed offset) in lockstep with the rest of the wave: one word for the free-path exponential, 16-bit fields for
// the sub-voxel jitter, one word for the collision decision. Scatter events draw from a second counter space, the
// camera jitter from a third, so a path's draws never depend on the lane schedule.
// ------------------------------------------------------------------------------------------------------------
constexpr uint32_t kEventSpace = 0x80000000u;
constexpr uint32_t kCameraBlock = 0xFFFFFFFFu;

enum : uint32_t {
    PF_ALIVE = 1u,
    PF_EXHAUSTED = 2u,
    PF_HAS_VALUE = 4u,
    PF_SHADOW = 8u,
    PF_FIRST = 16u,
    PF_PENDING = 32u,   // the lane waits for the wave to resolve its segment end or real collision together
    PF_PEND_END = 64u,  // the pending event is a segment end (else a real scattering collision)
};
constexpr uint32_t kFeatureBlock = 0xFFFFFFFEu;  // Philox block of the first event's finite-difference jitter
// flags bits 8..31 count the scattering events of the main path (in-scattering iterations for residual ratio
// tracking)
constexpr uint32_t kBounceShift = 8;

struct PathState {
    uint32_t unit;     // launch-local path index
    uint32_t pix;      // pixel index
    uint32_t frame;    // global frame (sample) number
    uint32_t step;     // next per-step Philox block
    uint32_t events;   // next event Philox block
    uint32_t flags;
    uint32_t taps;     // density lookups of this lane since the launch started (work counter)
    float thr;         // ratio weight; shadow-ray transmittance while PF_SHADOW; running T of residual ratio
    V3 color;          // next-event accumulated radiance; residual ratio tracking's accumulated radiance
    V3 shadowScale;    // radiance factor of the pending shadow ray
    V3 mainX, mainW;   // where the main path resumes after the shadow ray
    float wSum, resT, resDist;  // residual ratio tracking reservoir
    V3 weights;        // spectral tracking path weights
    float pendRho;     // density at the pending real collision
};

CRVPT_DEV uint4 pathBlock(const SceneParams& P, const PathState& st, uint32_t c0) {
    return philox4x32_10(make_uint4(c0, st.pix, st.frame, P.seedOffset), make_uint2(0x6A09E667u, 0x3C6EF372u));
}
CRVPT_DEV float u01(uint32_t x) { return float(x >> 8) * (1.0f / 16777216.0f); }
CRVPT_DEV float uOpen(uint32_t x) { return float((x >> 8) + 1u) * (1.0f / 16777216.0f); }
CRVPT_DEV float jitterHi(uint32_t x) { return (float(x >> 16) + 0.5f) * (1.0f / 65536.0f) - 0.5f; }
CRVPT_DEV float jitterLo(uint32_t x) { return (float(x & 0xFFFFu) + 0.5f) * (1.0f / 65536.0f) - 0.5f; }

// A finished path leaves its radiance in st.color and whether it met the medium in PF_HAS_VALUE; the kernel that
// runs it stores the result (one record per path, or a sum over the samples of a batch unit).
CRVPT_DEV void finishPath(PathState& st, V3 c, bool hasValue) {
    st.color = c;
    st.flags = hasValue ? PF_HAS_VALUE : 0u;
}
CRVPT_DEV void writePath(float4* out, const PathState& st) {
    out[st.unit] = make_float4(st.color.x, st.color.y, st.color.z, (st.flags & PF_HAS_VALUE) ? 1.0f : 0.0f);
}

// Uniform direction on the sphere from two uniforms, in the basis of d (randomDirection).
CRVPT_DEV V3 sphereDirection(float r1, float r2u, V3 d) {
    const float r2 = r2u * 2.0f - 1.0f;
    const float s = sqrtf(fmaxf(0.0f, 1.0f - r2 * r2));
    const float phi = kTwoPi * r1;
    V3 t0, t1;
    orthonormalBasis(d, t0, t1);
    return t0 * (__cosf(phi) * s) + t1 * (__sinf(phi) * s) + d * r2;
}

// importanceSamplePhase from two uniforms (HG inversion; isotropic below |g| < 0.001), with the scene's phase
// constants.
CRVPT_DEV V3 phaseDirection(const SceneParams& P, V3 d, float u0, float u1) {
    if (fabsf(P.phaseG) < 0.001f) return sphereDirection(u0, u1, -d);
    const float phi = u0 * kTwoPi;
    const float t = P.phaseC[PHASE_ONE_MINUS_G2] * rcpf(P.phaseC[PHASE_TWO_G] * u1 + P.phaseC[PHASE_ONE_MINUS_G]);
    const float cosTheta = P.phaseC[PHASE_HALF_INV_G] * (P.phaseC[PHASE_ONE_PLUS_G2] - t * t);
    const float sinTheta = sqrtf(fmaxf(0.0f, 1.0f - cosTheta * cosTheta));
    V3 t0, t1;
    orthonormalBasis(d, t0, t1);
    return t0 * (sinTheta * __sinf(phi)) + t1 * (sinTheta * __cosf(phi)) + d * cosTheta;
}

// Where the path kernels read the camera matrix and the sky: the device copy of the constants (P.mem) keeps them out
// of the scalar registers. Measured on the 512^3 sparse sweep: the camera from memory makes delta, ratio and
// decomposition tracking 7-9 % faster and next-event tracking 3.5 % slower; the sky from memory makes residual ratio
// tracking 8 % faster and delta and ratio tracking 6-7 % slower; the refill's claim and pixel arithmetic
// (launchParams: image size, frame range, divisors) from memory: 1.2 % over the whole sweep, next-event tracking
// unchanged (it keeps the launch constants).
template <int MODE>
CRVPT_DEV const float* poolCamera(const SceneParams& P) {
    return isNee<MODE>() ? P.invViewProj : P.mem->invViewProj;
}
template <int MODE>
CRVPT_DEV const SceneParams& launchParams(const SceneParams& P) {
    return isNee<MODE>() ? P : *P.mem;
}
template <int MODE>
CRVPT_DEV const SceneParams& skyParams(const SceneParams& P) {
    return (MODE == VPT_RESIDUAL_RATIO_TRACKING || MODE == VPT_DECOMPOSITION_TRACKING) ? *P.mem : P;
}

// Camera ray of the path: pixel jitter from the camera block (Clouds.glsl:80-90).
CRVPT_DEV void cameraRay(const SceneParams& P, const float* cam, const PathState& st, V3& x, V3& w) {
    const uint4 b = pathBlock(P, st, kCameraBlock);
    const uint32_t py = udiv31(st.pix, P.divWidth), px = st.pix - py * uint32_t(P.width);
    const float sx = 2.0f * (float(px) + u01(b.x)) / P.widthF - 1.0f;
    const float sy = 2.0f * (float(py) + u01(b.y)) / P.heightF - 1.0f;
    createCameraRay(cam, sx, sy, x, w);
}

// Starts the camera path of (st.pix, st.frame) through the inverse view-projection `cam`. False when the camera ray
// misses the box (the path is finished).
template <int MODE>
CRVPT_DEV bool startPathAt(PathState& st, Walker& W, const SceneParams& P, const SupervoxelGrid& S, float4* feat,
                           const float* cam) {
    st.step = 0;
    st.events = 0;
    st.flags = PF_ALIVE;
    st.thr = 1.0f;
    st.color = v3(0.0f);
    st.wSum = 0.0f;
    st.resT = 0.0f;
    st.resDist = 0.0f;
    st.weights = v3(1.0f);
    if (feat) feat[3 * size_t(st.unit)] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);  // no first event yet
    V3 x, w;
    cameraRay(P, cam, st, x, w);
    if (!walkerStart<MODE>(W, P, S, x, w, kInf)) {
        finishPath(st, skyRadiance(skyParams<MODE>(P), w), false);
        return false;
    }
    return true;
}

// Starts the path of launch unit st.unit = sample * (W*H) + pixel.
template <int MODE>
CRVPT_DEV bool startPath(PathState& st, Walker& W, const SceneParams& P, const SupervoxelGrid& S, float4* feat) {
    const SceneParams& Q = launchParams<MODE>(P);
    const uint32_t s = udiv31(st.unit, Q.divPixels);
    st.pix = st.unit - s * (uint32_t(Q.width) * uint32_t(Q.height));
    st.frame = Q.frameBase + s;
    return startPathAt<MODE>(st, W, P, S, feat, poolCamera<MODE>(P));
}

// getCloudFiniteDifference (VptUtils.glsl:726-748) at p for the normal feature map: central differences of the
// normalised grid one texel apart around the (stochastically jittered) texel, scaled by 100 / dim, normalised; the
// reference returns 0 for NanoVDB grids.
template <bool WIDE, bool TRI>
CRVPT_DEV V3 cloudFiniteDifference(const NanoAccessT<WIDE, TRI>&, const SceneParams&, const PathState&, V3) {
    return v3(0.0f);
}
template <bool TRI>
CRVPT_DEV V3 cloudFiniteDifference(const DenseAccessT<TRI>& grid, const SceneParams& P, const PathState& st, V3 p) {
    V3 u = grid.toTex(P, p);
    if (P.interpolation == INTERP_STOCHASTIC) {
        const uint4 b = pathBlock(P, st, kFeatureBlock);
        u = u + v3(u01(b.x) - 0.5f, u01(b.y) - 0.5f, u01(b.z) - 0.5f);
    }
    const V3 dim = v3(P.worldToTex[0] * (P.boxMax[0] - P.boxMin[0]), P.worldToTex[1] * (P.boxMax[1] - P.boxMin[1]),
                      P.worldToTex[2] * (P.boxMax[2] - P.boxMin[2]));
    const float gx = grid.lookupTex(P, u - v3(1.0f, 0.0f, 0.0f), 0.0f, 0.0f, 0.0f) -
                     grid.lookupTex(P, u + v3(1.0f, 0.0f, 0.0f), 0.0f, 0.0f, 0.0f);
    const float gy = grid.lookupTex(P, u - v3(0.0f, 1.0f, 0.0f), 0.0f, 0.0f, 0.0f) -
                     grid.lookupTex(P, u + v3(0.0f, 1.0f, 0.0f), 0.0f, 0.0f, 0.0f);
    const float gz = grid.lookupTex(P, u - v3(0.0f, 0.0f, 1.0f), 0.0f, 0.0f, 0.0f) -
                     grid.lookupTex(P, u + v3(0.0f, 0.0f, 1.0f), 0.0f, 0.0f, 0.0f);
    V3 d = v3(gx / dim.x * 100.0f, gy / dim.y * 100.0f, gz / dim.z * 100.0f);
    const float l = length(d);
    if (l > 1e-5f) d = d * (1.0f / l);
    return d;
}

// First scattering or absorption event of the path (the ScatterEvent the trackers fill, Clouds.glsl:91-96): its
// position, depth (distance along the camera ray), outgoing direction and phase pdf, extinction-scaled density and
// the finite-difference normal, written once per path for the feature maps.
template <int MODE, class Grid>
CRVPT_DEV void recordFirst(PathState& st, const SceneParams& P, const Grid& grid, float4* feat, V3 p, V3 w,
                           float pdfW, float dens, float depth) {
    if (!feat || (st.flags & PF_FIRST)) return;
    st.flags |= PF_FIRST;
    const V3 n = cloudFiniteDifference(grid, P, st, p);
    const size_t k = 3 * size_t(st.unit);
    feat[k] = make_float4(p.x, p.y, p.z, depth);
    feat[k + 1] = make_float4(w.x, w.y, w.z, pdfW);
    feat[k + 2] = make_float4(n.x, n.y, n.z, dens);
}

// Real scattering collision at p of the main path (rare compared with tentative collisions).
template <int MODE, class Grid>
CRVPT_DEV bool scatterEvent(PathState& st, Walker& W, const SceneParams& P, const SupervoxelGrid& S, V3 p,
                            float pAbsOverReal, float rho, const Grid& grid, float4* feat) {
    const uint4 e = pathBlock(P, st, kEventSpace | st.events++);
    st.flags += 1u << kBounceShift;
    const V3 nw = phaseDirection(P, W.w, u01(e.x), u01(e.y));
    if (feat && MODE != VPT_DECOMPOSITION_TRACKING) {
        // ratio tracking records neither depth nor density (RatioTracking.glsl:75-81)
        const bool ratio = MODE == VPT_RATIO_TRACKING;
        const float ext = (isNee<MODE>() || isSpectral<MODE>()) ? P.extinctionMax : P.extinction;
        recordFirst<MODE>(st, P, grid, feat, p, nw, evalPhase(P, W.w, nw), ratio ? 0.0f : rho * ext,
                          ratio ? 0.0f : W.t);
    }
    if (MODE == VPT_RATIO_TRACKING) st.thr *= 1.0f - pAbsOverReal;
    st.flags |= PF_HAS_VALUE;
    if (isNee<MODE>()) {
        bool toLight = false;
        float probFactor = 1.0f, dist = 0.0f;
        int li = 0;
        V3 dir;
        if (P.numLights > 0) {
            toLight = P.isEnvMapBlack ? true : (u01(e.z) > 0.5f);
            probFactor = P.isEnvMapBlack ? 1.0f : 2.0f;
        }
        if (toLight) {
            if (P.numLights > 1) {
                li = min(int(u01(e.w) * float(P.numLights)), P.numLights - 1);
                probFactor *= float(P.numLights);
            }
            const Light& L = P.lights[li];
            const V3 lp = lightPositionWorld(P, L);
            const V3 d = (L.lightType == 0u || L.lightType == 1u) ? lp - p : lp;
            dist = length(d);
            dir = d * rcpf(dist + 1e-5f);
        } else {
            const uint4 e2 = pathBlock(P, st, kEventSpace | st.events++);
            dir = sphereDirection(u01(e2.x), u01(e2.y), v3(1.0f, 0.0f, 0.0f));
        }
        const float ph = evalPhase(P, W.w, dir);
        V3 L = toLight ? lightRadiance(P, p, li) * (probFactor * ph)
                       : skyRadiance(skyParams<MODE>(P), dir) * (probFactor * ph * (4.0f * kPi));
        if (isSpectral<MODE>()) L = L * vmin(st.weights, 100000.0f);
        st.mainX = p;
        st.mainW = nw;
        if ((L.x > 0.0f || L.y > 0.0f || L.z > 0.0f) &&
            walkerStart<MODE>(W, P, S, p, dir, toLight ? dist : kInf, true)) {
            st.flags |= PF_SHADOW;
            st.shadowScale = L;
            st.thr = 1.0f;
            return true;
        }
        st.color += L;  // zero contribution, or the shadow segment lies outside the box (T = 1)
    }
    if (MODE == VPT_DECOMPOSITION_TRACKING) {
        // a real collision absorbs with probability 1 - albedo (DecompositionTracking.glsl:92-98)
        if (u01(e.z) < P.absorption) {
            finishPath(st, v3(0.0f), true);
            return false;
        }
    }
    if (!walkerStart<MODE>(W, P, S, p, nw, kInf, false, MODE == VPT_DECOMPOSITION_TRACKING)) {
        const V3 c = isNee<MODE>() ? st.color
                                   : skyRadiance(skyParams<MODE>(P), nw) *
                                         (isSpectral<MODE>() ? vmin(st.weights, 100000.0f) : v3(st.thr));
        finishPath(st, c, true);
        return false;
    }
    return true;
}

// End of the current segment: a finished shadow ray resumes its main path, a main path leaves the volume.
template <int MODE>
CRVPT_DEV bool endEvent(PathState& st, Walker& W, const SceneParams& P, const SupervoxelGrid& S) {
    if (isNee<MODE>() && (st.flags & PF_SHADOW)) {
        st.color += st.shadowScale * st.thr;
        st.flags &= ~PF_SHADOW;
        if (walkerStart<MODE>(W, P, S, st.mainX, st.mainW, kInf)) return true;
        finishPath(st, st.color, true);
        return false;
    }
    const bool hv = (st.flags & PF_HAS_VALUE) != 0;
    V3 c;
    const V3 wts = isSpectral<MODE>() ? vmin(st.weights, 100000.0f) : v3(st.thr);
    if (isNee<MODE>()) c = hv ? st.color : st.color + skyRadiance(skyParams<MODE>(P), W.w) * wts;
    else c = skyRadiance(skyParams<MODE>(P), W.w) * wts;
    finishPath(st, c, hv);
    return false;
}

// Residual ratio tracking: end of one traversal of the box (ResidualRatioTracking.glsl:196-219). The traversal's
// sky term is added; the reservoir sample becomes the next in-scattering point unless xi > wSum or ten iterations
// were made.
CRVPT_DEV bool rrtTraversalEnd(PathState& st, Walker& W, const SceneParams& P, const SupervoxelGrid& S) {
    constexpr int MODE = VPT_RESIDUAL_RATIO_TRACKING;
    const uint4 e = pathBlock(P, st, kEventSpace | st.events++);
    st.color += skyRadiance(skyParams<MODE>(P), W.w) * st.thr;
    const uint32_t iteration = st.flags >> kBounceShift;
    if (u01(e.x) > st.wSum || iteration >= 10u) {
        finishPath(st, st.color, true);
        return false;
    }
    st.flags += 1u << kBounceShift;
    const V3 x = W.o + W.w * st.resDist;
    const V3 nw = phaseDirection(P, W.w, u01(e.y), u01(e.z));
    st.thr = st.resT;
    st.wSum = 0.0f;
    st.resT = 0.0f;
    st.resDist = 0.0f;
    if (!walkerStart<MODE>(W, P, S, x, nw, kInf)) {
        st.color += skyRadiance(skyParams<MODE>(P), nw) * st.thr;
        finishPath(st, st.color, true);
        return false;
    }
    return true;
}

// Tentative collision of a spectral tracker's main path at p (DeltaTracking.glsl:87-160,
// NextEventTracking.glsl:89-215): collision probabilities from the path-history-weighted averages of the RGB
// absorption, scattering and null coefficients; the weights carry the ratio of each channel's coefficient to the
// chosen probability and are clamped to 100.
template <int MODE, class Grid>
CRVPT_DEV bool spectralCollision(PathState& st, Walker& W, const SceneParams& P, const SupervoxelGrid& S, V3 p,
                                 float rho, float xi, const Grid& grid, float4* feat) {
    const V3 ext = v3(P.extinctionRGB) * rho;
    const V3 alb = v3(P.albedoRGB);
    const V3 sa = (v3(1.0f) - alb) * ext;
    const V3 ss = alb * ext;
    const V3 sn = v3(W.mu) - ext;
    float Pa = avg3(sa * st.weights), Ps = avg3(ss * st.weights), Pn = avg3(sn * st.weights);
    const float C = Pa + Ps + Pn;
    Pa /= C;
    Ps /= C;
    Pn /= C;
    if (xi < Pa) {
        recordFirst<MODE>(st, P, grid, feat, p, v3(0.0f), 0.0f, rho * P.extinctionMax, W.t);
        finishPath(st, isNee<MODE>() ? st.color : v3(0.0f), true);
        return false;
    }
    if (xi < Pa + Ps) {
        if (P.maxBounces != 0u && (st.flags >> kBounceShift) >= P.maxBounces) {
            finishPath(st, isNee<MODE>() ? st.color : v3(0.0f), true);
            return false;
        }
        st.weights = st.weights * (ss / v3(W.mu * Ps));
        // next-event estimation uses the weights before the clamp, as the reference's loop clamps at its end
        if (!isNee<MODE>()) st.weights = vmin(st.weights, 100.0f);
        const bool alive = scatterEvent<MODE>(st, W, P, S, p, 0.0f, rho, grid, feat);
        st.weights = vmin(st.weights, 100.0f);
        return alive;
    }
    st.weights = vmin(st.weights * (sn / v3(W.mu * Pn)), 100.0f);
    return true;
}

// One step of a live path. Returns false when the path finished (its radiance is written).
template <int MODE, class Grid>
CRVPT_DEV bool stepPath(PathState& st, Walker& W, const SceneParams& P, const SupervoxelGrid& S, const Grid& grid,
                        typename Grid::Cache& gc, float4* feat) {
    const uint4 r = pathBlock(P, st, st.step++);
    // Exp(1) free-path budget -ln(u) = -log2(u) ln 2 from the hardware log2 (u in [2^-24, 1]: no denormal inputs)
    const int ev = walkAdvance<MODE>(W, P, S, __builtin_amdgcn_logf(uOpen(r.x)) * -0.693147182f);
    if (ev == WALK_MOVED) return true;
    if (MODE == VPT_RESIDUAL_RATIO_TRACKING) {
        st.thr *= __expf(-W.ctrl);  // control transmittance of the walked stretch
        W.ctrl = 0.0f;
        if (ev == WALK_END) {
            st.flags |= PF_PENDING | PF_PEND_END;
            return true;
        }
        const float rho = grid.lookupWorldJ(P, W.o + W.w * W.t, jitterHi(r.y), jitterLo(r.y), jitterHi(r.z), gc);
        st.taps++;
        st.thr *= 1.0f - (P.extinction * rho - W.muC) * rcpf(W.mu);
        st.wSum += st.thr;
        if (u01(r.w) * st.wSum < st.thr) {
            st.resT = st.thr;
            st.resDist = W.t;
        }
        return true;
    }
    if (ev == WALK_END) {
        if (isSpectral<MODE>()) return endEvent<MODE>(st, W, P, S);
        st.flags |= PF_PENDING | PF_PEND_END;
        return true;
    }
    const V3 p = W.o + W.w * W.t;
    const float xi = u01(r.w);
    if (MODE == VPT_DECOMPOSITION_TRACKING) {
        // control collision (xi * mu < mu_c) needs no lookup; a residual one is real with prob (sigma rho - mu_c)/mu_r
        const float xm = xi * W.mu;
        if (!(xm < W.muC)) {
            const float rho = grid.lookupWorldJ(P, p, jitterHi(r.y), jitterLo(r.y), jitterHi(r.z), gc);
            st.taps++;
            if (!(xm < P.extinction * rho)) return true;
        }
        if (P.maxBounces != 0u && (st.flags >> kBounceShift) >= P.maxBounces) {
            finishPath(st, v3(0.0f), true);
            return false;
        }
        st.pendRho = 0.0f;
        st.flags |= PF_PENDING;
        return true;
    }
    const float rho = grid.lookupWorldJ(P, p, jitterHi(r.y), jitterLo(r.y), jitterHi(r.z), gc);
    st.taps++;
    const float pReal = P.extinction * rho * rcpf(W.mu);
    if (isSpectral<MODE>() && !(st.flags & PF_SHADOW))
        return spectralCollision<MODE>(st, W, P, S, p, rho, xi, grid, feat);
    if (isNee<MODE>() && (st.flags & PF_SHADOW)) {
        st.thr *= 1.0f - pReal;  // ratio tracking (NeeTransmittance.glsl:107-113)
        // Once the product is exactly 0 every later factor leaves it 0 and the shadow ray adds exactly 0.
        if (st.thr == 0.0f) st.flags |= PF_PENDING | PF_PEND_END;
        return true;
    }
    if (!(xi < pReal)) return true;  // null collision
    const float pAbs = P.absorption * pReal;
    const bool bounceLimit = P.maxBounces != 0u && (st.flags >> kBounceShift) >= P.maxBounces;
    if ((MODE != VPT_RATIO_TRACKING && xi < pAbs) || bounceLimit) {
        // absorption or a scattering event beyond the bounce limit: delta / ratio end with 0, next-event with its sum
        if (MODE != VPT_RATIO_TRACKING && xi < pAbs) {
            // delta tracking records w = 0 (DeltaTracking.glsl:291-300), next-event the incoming direction
            recordFirst<MODE>(st, P, grid, feat, p, isNee<MODE>() ? W.w : v3(0.0f), 0.0f,
                              rho * (isNee<MODE>() ? P.extinctionMax : P.extinction), W.t);
        }
        finishPath(st, isNee<MODE>() ? st.color : v3(0.0f), true);
        return false;
    }
    st.pendRho = rho;
    st.flags |= PF_PENDING;
    return true;
}

// Resolves a lane's pending segment end or real scattering collision (the walker still stands at the event).
// Deferring these rarer, longer branches lets a wave run them once for many lanes instead of once per step for
// whichever lane happens to need them; the path's random numbers are counter-based, so the result is the same.
template <int MODE, class Grid>
CRVPT_DEV bool resolvePending(PathState& st, Walker& W, const SceneParams& P, const SupervoxelGrid& S,
                              const Grid& grid, float4* feat) {
    const bool end = (st.flags & PF_PEND_END) != 0;
    st.flags &= ~(PF_PENDING | PF_PEND_END);
    if (end) {
        if (MODE == VPT_RESIDUAL_RATIO_TRACKING) return rrtTraversalEnd(st, W, P, S);
        return endEvent<MODE>(st, W, P, S);
    }
    const V3 p = W.o + W.w * W.t;
    // analog ratio tracking weights a real collision by its scattering probability, the albedo
    return scatterEvent<MODE>(st, W, P, S, p, MODE == VPT_RATIO_TRACKING ? P.absorption : 0.0f, st.pendRho,
                              grid, feat);
}

// Wave-reduced 64-bit work counters: [0] density taps, [1] supervoxel majorant / table reads, [2] paths started,
// [3] sparse-tree descents (lookup memo misses).
CRVPT_DEV void flushCounters(unsigned long long* stats, uint32_t taps, uint32_t svs, uint32_t paths,
                             uint32_t descents) {
    unsigned long long v[4] = {taps, svs, paths, descents};
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] += __shfl_down(v[q], off);
    if (__lane_id() == 0)
#pragma unroll
        for (int q = 0; q < 4; q++) atomicAdd(&stats[q], v[q]);
}

#ifndef CRVPT_RESOLVE_LANES
#define CRVPT_RESOLVE_LANES 24
#endif
// Pending lanes that trigger a resolve pass of the wave (512^3 sparse sweep of 4 estimators x 3 g at 1024^2 x 256 spp,
// 8 / 16 / 24 / 32 lanes: 1546 / 1652 / 1669 / 1650 Mpaths/s).
constexpr int kResolveLanes = CRVPT_RESOLVE_LANES;

}  // namespace crvpt
