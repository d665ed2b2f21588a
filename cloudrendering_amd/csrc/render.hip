// gfx950 kernels of the volumetric path tracer.
//
// Work decomposition: a "unit" is one camera path (pixel p, sample s) with record slot s * (W*H) + p. Claims walk
// the image in 16 x 16 screen tiles, all of a launch's samples of one tile before the next, so a wave's 64
// consecutive claims are 64 neighbouring pixels of one sample (coherent primary rays) and the paths in flight read a
// small region of the volume. A persistent grid of waves drains a global claim counter: whenever lanes of a wave have no live path, one lane claims as many units as the
// wave has free lanes (ballot + popcount + one atomic) and the free lanes take them in lane order (mbcnt), so
// lanes refill at any step instead of idling until the longest path of the wave ends. Each path writes its
// radiance to its own slot; a separate kernel sums the slots of a pixel in sample order, so images are
// bit-reproducible whatever the lane/wave schedule was.
//
// Free-flight sampling walks a supervoxel majorant grid (piecewise-constant majorant along the ray): an Exp(1)
// optical budget is spent supervoxel by supervoxel, empty supervoxels cost one DDA step and no density fetch, and
// a tentative collision is accepted with probability sigma*rho / mu_local. With the grid disabled the walker
// degenerates to the reference's global majorant (extinction.x), i.e. the reference trackers' exact free-path
// distribution.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>

#include "path_tracer.hpp"

namespace crvpt {

#ifndef CRVPT_CLAIM_TILE
#define CRVPT_CLAIM_TILE 16
#endif
// Claims walk the image in square screen tiles of this edge, all of a launch's samples of one tile before the next
// (0: sample-major over the whole image). The paths in flight then cover a few tiles instead of a quarter of the
// image, so the volume region they read is small enough to stay in the L2 / Infinity Cache.
constexpr uint32_t kClaimTile = CRVPT_CLAIM_TILE;

CRVPT_DEV uint32_t claimCount(const SceneParams& P) {
    if (kClaimTile == 0u) return uint32_t(P.numUnits);
    const uint32_t tx = (uint32_t(P.width) + kClaimTile - 1u) / kClaimTile;
    const uint32_t ty = (uint32_t(P.height) + kClaimTile - 1u) / kClaimTile;
    return tx * ty * kClaimTile * kClaimTile * P.numFrames;
}

// Claim index -> launch slot sample * W*H + pixel (the record index); false for padding claims outside the image.
CRVPT_DEV bool claimToSlot(const SceneParams& P, uint32_t c, uint32_t& slot) {
    if (kClaimTile == 0u) {
        slot = c;
        return true;
    }
    constexpr uint32_t tp = kClaimTile * kClaimTile;
    const uint32_t w = uint32_t(P.width), h = uint32_t(P.height);
    const uint32_t tx = (w + kClaimTile - 1u) / kClaimTile;
    const uint32_t perTile = tp * P.numFrames;
    const uint32_t tile = udiv31(c, P.divTileClaims), rem = c - tile * perTile;
    const uint32_t s = rem / tp, q = rem - s * tp;
    const uint32_t ty = udiv31(tile, P.divTilesX);
    const uint32_t x = (tile - ty * tx) * kClaimTile + q % kClaimTile, y = ty * kClaimTile + q / kClaimTile;
    if (x >= w || y >= h) return false;
    slot = s * (w * h) + y * w + x;
    return true;
}

#ifndef CRVPT_CLAIM_PARTS
#define CRVPT_CLAIM_PARTS 32
#endif
// Claim partitions: the tiles are split into this many contiguous ranges, each with its own claim counter, and a
// workgroup first drains range blockIdx.x % parts, then the following ones. Workgroups are dispatched to the 8 XCDs
// round robin, so with a multiple of 8 ranges the paths in flight on one XCD come from a few bands of the image,
// which keeps the volume region they read, and so the XCD's L2 working set, small; the claims also spread over many
// counters instead of one.
constexpr uint32_t kClaimParts = CRVPT_CLAIM_PARTS;
#ifndef CRVPT_COUNTER_STRIDE
#define CRVPT_COUNTER_STRIDE 32
#endif
// Distance in 32-bit words between the claim counters of consecutive ranges (32: one 128-byte line each).
constexpr uint32_t kCounterStride = CRVPT_COUNTER_STRIDE;

CRVPT_DEV uint32_t partBegin(const SceneParams& P, uint32_t part) {
    if (kClaimParts == 1u) return part == 0u ? 0u : claimCount(P);
    const uint32_t tx = (uint32_t(P.width) + kClaimTile - 1u) / kClaimTile;
    const uint32_t ty = (uint32_t(P.height) + kClaimTile - 1u) / kClaimTile;
    const uint32_t tiles = tx * ty;
    return uint32_t(uint64_t(tiles) * part / kClaimParts) * (kClaimTile * kClaimTile * P.numFrames);
}

// STATS: instrumented variant that keeps the per-lane work counters and adds them to `stats` at exit; in the timed
// variant the counters are dead code and take no registers.
template <int MODE, class Grid, bool STATS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(poolWaves<MODE>()))) void k_pathPool(
        SceneParams P, Grid grid, SupervoxelGrid S, float4* __restrict__ out, uint32_t* __restrict__ counter,
        int stepsPerRefill, unsigned long long* __restrict__ stats, float4* __restrict__ feat) {
    const int lane = __lane_id();
    PathState st;
    st.flags = 0;
    st.unit = 0;
    st.taps = 0;
    bool exhausted = false;
    uint32_t paths = 0;
    Walker W;
    W.svReads = 0;
    typename Grid::Cache gc;
    uint32_t part = kClaimParts == 1u ? 0u : blockIdx.x % kClaimParts;  // wave-uniform
    uint32_t partsLeft = kClaimParts;
    uint32_t begin = partBegin(P, part), count = partBegin(P, part + 1u) - begin;
    while (true) {
        // Refill until every lane holds a live path or the queue is drained: a claimed path whose camera ray misses
        // the volume finishes inside startPath, and its lane claims again before the wave steps.
        while (true) {
            const bool needs = !(st.flags & PF_ALIVE) && !exhausted;
            const uint64_t needMask = __ballot(needs);
            if (needMask == 0) break;
            const int leader = __ffsll((unsigned long long)needMask) - 1;
            const uint32_t want = uint32_t(__popcll(needMask));
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&counter[part * kCounterStride], want);
            base = __builtin_amdgcn_readfirstlane(__shfl(base, leader));  // wave-uniform: the range state in SGPRs
            if (needs) {
                const uint32_t c = base + uint32_t(__popcll(needMask & ((1ull << lane) - 1ull)));
                if (c < count && claimToSlot(launchParams<MODE>(P), begin + c, st.unit)) {
                    paths++;
                    if (!startPath<MODE>(st, W, P, S, feat)) writePath(out, st);
                }
            }
            if (base + want >= count) {
                // this range is drained: the next one, until every range is
                if (--partsLeft == 0u) {
                    exhausted = true;
                } else {
                    part = part + 1u == kClaimParts ? 0u : part + 1u;
                    begin = partBegin(P, part);
                    count = partBegin(P, part + 1u) - begin;
                }
            }
        }
        if (__ballot((st.flags & PF_ALIVE) != 0) == 0) break;  // every lane is exhausted
        for (int k = 0; k < stepsPerRefill; k++) {
            if ((st.flags & (PF_ALIVE | PF_PENDING)) == PF_ALIVE) {
                if (!stepPath<MODE, Grid>(st, W, P, S, grid, gc, feat)) {
                    writePath(out, st);
                    st.flags = 0;
                }
            }
            const uint64_t pend = __ballot((st.flags & PF_PENDING) != 0);
            if (pend != 0ull && (__popcll(pend) >= kResolveLanes || k == stepsPerRefill - 1)) {
                if (st.flags & PF_PENDING) {
                    if (!resolvePending<MODE, Grid>(st, W, P, S, grid, feat)) {
                        writePath(out, st);
                        st.flags = 0;
                    }
                }
            }
        }
    }
    if (STATS) flushCounters(stats, st.taps, W.svReads, paths, gc.descents());
}

// ------------------------------------------------------------------------------------------------------------
// Helpers of the batched residual-ratio and decomposition transmittance queries (reference traversal order).
// ------------------------------------------------------------------------------------------------------------
// Per-lane work counters of the transmittance estimators (density taps, supervoxel table reads).
struct WorkCount {
    uint32_t taps = 0;
    uint32_t svReads = 0;
};

CRVPT_DEV float signf(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }
CRVPT_DEV float fractf(float v) { return v - floorf(v); }

CRVPT_DEV bool refInside(const SupervoxelGrid& S, int x, int y, int z) {
    return x >= 0 && y >= 0 && z >= 0 && x < S.refGrid[0] && y < S.refGrid[1] && z < S.refGrid[2];
}
CRVPT_DEV size_t refIndex(const SupervoxelGrid& S, int x, int y, int z) {
    return (size_t(z) * S.refGrid[1] + y) * S.refGrid[0] + x;
}
CRVPT_DEV V3 gridDims(const SceneParams& P) {
    return v3(P.worldToTex[0] * (P.boxMax[0] - P.boxMin[0]), P.worldToTex[1] * (P.boxMax[1] - P.boxMin[1]),
              P.worldToTex[2] * (P.boxMax[2] - P.boxMin[2]));
}

struct RefDda {
    int step[3];
    float tMax[3], tDelta[3];
    int idx[3];
};

CRVPT_DEV bool refDdaInit(const SceneParams& P, const SupervoxelGrid& S, V3 x, V3 w, float dTotal, RefDda& q) {
    const V3 dims = gridDims(P);
    const V3 boxDelta = v3(P.boxMax) - v3(P.boxMin);
    const V3 sp = (x - v3(P.boxMin)) / boxDelta * dims / v3(float(S.refSize));
    const V3 ep = (x + w * dTotal - v3(P.boxMin)) / boxDelta * dims / v3(float(S.refSize));
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float s0 = comp(sp, a), e0 = comp(ep, a);
        const int st = int(signf(e0 - s0));
        const float td = st != 0 ? fminf(float(st) / (e0 - s0), 1e7f) : 1e7f;
        q.step[a] = st;
        q.tDelta[a] = td;
        q.tMax[a] = st > 0 ? td * (1.0f - fractf(s0)) : td * fractf(s0);
        q.idx[a] = int(floorf(s0));
    }
    return !(q.step[0] == 0 && q.step[1] == 0 && q.step[2] == 0);
}

CRVPT_DEV void refDdaAdvance(RefDda& q) {
    int a;
    if (q.tMax[0] < q.tMax[1]) a = q.tMax[0] < q.tMax[2] ? 0 : 2;
    else a = q.tMax[1] < q.tMax[2] ? 1 : 2;
    q.idx[a] += q.step[a];
    q.tMax[a] += q.tDelta[a];
}

CRVPT_DEV void refSvBounds(const SceneParams& P, const SupervoxelGrid& S, const int idx[3], V3& mn, V3& mx) {
    const V3 dims = gridDims(P);
    const V3 boxDelta = v3(P.boxMax) - v3(P.boxMin);
    const V3 lo = v3(float(idx[0] * S.refSize), float(idx[1] * S.refSize), float(idx[2] * S.refSize));
    const V3 hi = lo + v3(float(S.refSize));
    mn = lo / dims * boxDelta + v3(P.boxMin);
    mx = hi / dims * boxDelta + v3(P.boxMin);
}

template <class Grid>
CRVPT_DEV float rrtEstimator(const SceneParams& P, const Grid& grid, Rng& rng, V3 x, V3 w, float dStart, float dEnd,
                             float T, float& wSum, float& resT, float& resDist, float muC, float muR,
                             WorkCount& wc, typename Grid::Cache& gc) {
    const float Tc = __expf(-muC * (dEnd - dStart));
    float Tr = 1.0f;
    float dist = dStart;
    while (true) {
        const float t = rng.exponential() / muR;
        x += w * t;
        dist += t;
        if (dist >= dEnd) break;
        const float rho = density(grid, P, rng, x, gc);
        wc.taps++;
        Tr *= 1.0f - (P.extinction * rho - muC) / muR;
        const float Tlocal = T * Tr * __expf(-muC * (dist - dStart));
        wSum += Tlocal;
        if (rng.uniform() < Tlocal / wSum) {
            resT = Tlocal;
            resDist = dist;
        }
    }
    return Tc * Tr;
}

// ------------------------------------------------------------------------------------------------------------
// Accumulation: running mean over all frames since the last reset (Clouds.glsl:172-179), summed per pixel in
// sample order, then written as [C, H, W] float32 (alpha = 1) when requested.
// ------------------------------------------------------------------------------------------------------------
// foreground: 0 plain running mean (alpha 1); otherwise the OUTPUT_FOREGROUND_MAP store of Clouds.glsl:181-191,
// the running mean of (radiance, 1) over samples whose first event has a value and (0, 0, 0, 0) over the others:
// 1 reads that flag from the record's w, 2 is the trackers that never record a first event (residual ratio and
// decomposition tracking), whose every sample stores 0.
__global__ void k_accumulate(const float4* __restrict__ paths, uint32_t npix, uint32_t frames,
                             float4* __restrict__ acc, uint32_t prevFrames, float* __restrict__ outCHW,
                             int channels, int foreground) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    // FP64 sums: a float32 sum of thousands of samples drifts systematically (tracking estimators return a few
    // distinct values, e.g. the sky radiance or 0, and adding the same value to a growing float sum rounds the same
    // way every time: about 5e-4 relative after 65536 samples); the doubles keep the mean exact to float precision.
    double sx = 0.0, sy = 0.0, sz = 0.0, sw = 0.0;
    for (uint32_t f = 0; f < frames; f++) {
        const float4 v = paths[size_t(f) * npix + p];
        const float keep = foreground == 0 ? 1.0f : (foreground == 1 ? v.w : 0.0f);
        sx += double(v.x * keep);
        sy += double(v.y * keep);
        sz += double(v.z * keep);
        sw += double(keep);
    }
    float4 a = prevFrames == 0 ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : acc[p];
    const double prev = double(prevFrames), inv = 1.0 / double(prevFrames + frames);
    a.x = float((double(a.x) * prev + sx) * inv);
    a.y = float((double(a.y) * prev + sy) * inv);
    a.z = float((double(a.z) * prev + sz) * inv);
    a.w = foreground == 0 ? 1.0f : float((double(a.w) * prev + sw) * inv);
    acc[p] = a;
    if (outCHW) {
        outCHW[p] = a.x;
        outCHW[npix + p] = a.y;
        outCHW[2 * size_t(npix) + p] = a.z;
        if (channels == 4) outCHW[3 * size_t(npix) + p] = a.w;
    }
}

__global__ void k_writeCHW(const float4* __restrict__ img, uint32_t npix, float* __restrict__ outCHW, int channels) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const float4 a = img[p];
    outCHW[p] = a.x;
    outCHW[npix + p] = a.y;
    outCHW[2 * size_t(npix) + p] = a.z;
    if (channels == 4) outCHW[3 * size_t(npix) + p] = a.w;
}

// Feature maps (Clouds.glsl:146-383): per pixel, the chunk's samples in order, each mixed into the running state
// with weight 1 / (frame + 1) as the reference's imageLoad / mix / imageStore sequence does.
__global__ void k_accumulateFeatures(SceneParams P, const float4* __restrict__ paths, const float4* __restrict__ feat,
                                     uint32_t npix, uint32_t frames, uint32_t prevFrames, int noAccumulation,
                                     FeatureAcc A) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const float px = float(p % uint32_t(P.width)), py = float(p / uint32_t(P.width));
    float4 firstX = A.firstX[p], firstW = A.firstW[p], normal = A.normal[p], cloudOnly = A.cloudOnly[p];
    float4 background = A.background[p], albedo = A.albedo[p], depth = A.depth[p], dens = A.dens[p];
    float4 flow = A.flow[p], flowRev = A.flowRev[p], radSq = A.radianceSq[p];
    const float* M = P.prevViewProj;
    for (uint32_t f = 0; f < frames; f++) {
        const uint32_t frame = prevFrames + f;
        const float a = 1.0f / float(frame + 1);
        const float b = 1.0f - a;
        const size_t u = size_t(f) * npix + p;
        const float4 res = paths[u];
        const float4 f0 = feat[3 * u], f1 = feat[3 * u + 1], f2 = feat[3 * u + 2];
        const bool has = f0.w >= 0.0f;
        if (frame == 0) {
            const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            firstX = normal = cloudOnly = background = albedo = depth = dens = radSq = z;
            flow = make_float4(0.0f, 0.0f, -1.0f, -1.0f);
        }
        PathState st;
        st.pix = p;
        st.frame = P.frameBase + f;
        V3 xo, w;
        cameraRay(P, P.invViewProj, st, xo, w);
        auto mix4 = [&](float4 o, float4 v) {
            return make_float4(o.x * b + v.x * a, o.y * b + v.y * a, o.z * b + v.z * a, o.w * b + v.w * a);
        };
        cloudOnly = mix4(cloudOnly, has ? make_float4(res.x, res.y, res.z, 1.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        radSq = mix4(radSq, make_float4(res.x * res.x, res.y * res.y, res.z * res.z, 0.0f));
        const V3 sky = sampleSky(P, w);
        background = mix4(background, has ? make_float4(sky.x, sky.y, sky.z, 1.0f) : make_float4(res.x, res.y, res.z, 1.0f));
        firstX = mix4(firstX, has ? make_float4(f0.x, f0.y, f0.z, 1.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        const float d = has ? f0.w : (noAccumulation ? P.farDistance : 0.0f);
        depth.x = depth.x * b + d * a;
        depth.y = depth.y * b + (has || !noAccumulation ? d * d : 0.0f) * a;
        depth.z = depth.z * b + (has ? f0.w : P.farDistance) * a;
        const float dn = has ? f2.w * 0.001f : 0.0f;
        dens.x = dens.x * b + dn * a;
        dens.y = dens.y * b + dn * dn * a;
        albedo = mix4(albedo, has ? make_float4(P.albedoRGB[0], P.albedoRGB[1], P.albedoRGB[2], 1.0f)
                                  : make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        float nx = 0.0f, ny = 0.0f;
        float2 fl = make_float2(0.0f, 0.0f);
        float2 uv = make_float2(flow.z, flow.w);
        if (has) {
            normal = make_float4(normal.x * b + f2.x * a, normal.y * b + f2.y * a, normal.z * b + f2.z * a, 1.0f);
            // inverseTransposedViewMatrix * (n, 0): rows of the inverse view's rotation columns
            const float* iv = P.invView;
            const float Acos = iv[0] * normal.x + iv[1] * normal.y + iv[2] * normal.z;
            const float Bcos = iv[4] * normal.x + iv[5] * normal.y + iv[6] * normal.z;
            nx = Acos / sqrtf(1.0f - Acos * Acos);
            ny = Bcos / sqrtf(1.0f - Bcos * Bcos);
            firstW = f1;
            const float cx = M[0] * f0.x + M[4] * f0.y + M[8] * f0.z + M[12];
            const float cy = M[1] * f0.x + M[5] * f0.y + M[9] * f0.z + M[13];
            const float cw = M[3] * f0.x + M[7] * f0.y + M[11] * f0.z + M[15];
            uv = make_float2(cx / cw * 0.5f + 0.5f, cy / cw * 0.5f + 0.5f);
            fl = make_float2(px - ((0.5f * (cx / cw) + 0.5f) * float(P.width) - 0.5f),
                             py - ((0.5f * (cy / cw) + 0.5f) * float(P.height) - 0.5f));
        } else {
            const float k = float(frame) / (float(frame) + 1.0f);
            normal = make_float4(normal.x * k, normal.y * k, normal.z * k, 1.0f);
            firstW = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        flow = make_float4(fl.x, fl.y, uv.x, uv.y);
        {
            const float wx = has ? f0.x : w.x, wy = has ? f0.y : w.y, wz = has ? f0.z : w.z, ww = has ? 1.0f : 0.0f;
            const float cx = M[0] * wx + M[4] * wy + M[8] * wz + M[12] * ww;
            const float cy = M[1] * wx + M[5] * wy + M[9] * wz + M[13] * ww;
            const float cw = M[3] * wx + M[7] * wy + M[11] * wz + M[15] * ww;
            flowRev = make_float4((0.5f * (cx / cw) + 0.5f) * float(P.width) - 0.5f - px,
                                  (0.5f * (cy / cw) + 0.5f) * float(P.height) - 0.5f - py, 0.0f, 0.0f);
        }
        if (frame != 0) {
            nx = dens.z * b + nx * a;
            ny = dens.w * b + ny * a;
        }
        dens.z = nx;
        dens.w = ny;
        float fw = fabsf(nx) + fabsf(ny);
        if (frame != 0) fw = depth.w * b + fw * a;
        depth.w = fw;
    }
    A.firstX[p] = firstX;
    A.firstW[p] = firstW;
    A.normal[p] = normal;
    A.cloudOnly[p] = cloudOnly;
    A.background[p] = background;
    A.albedo[p] = albedo;
    A.depth[p] = depth;
    A.dens[p] = dens;
    A.flow[p] = flow;
    A.flowRev[p] = flowRev;
    A.radianceSq[p] = radSq;
}

// FeatureMapTypeVpt (VolumetricPathTracingPass.hpp:67-73) blitted to RGBA as the reference's readback does: RG and R
// formats fill the missing channels with 0 and alpha with 1.
__global__ void k_featureOut(int map, FeatureAcc A, const float4* __restrict__ result, uint32_t npix,
                             float* __restrict__ outCHW, int channels) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    float4 v;
    switch (map) {
        case 1: v = A.firstX[p]; break;
        case 2: v = A.firstW[p]; break;
        case 3: v = A.normal[p]; break;
        case 4: {  // NormalizeNormals.glsl: unit length where the length exceeds 1e-5
            const float4 n = A.normal[p];
            const float l = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
            v = l > 1e-5f ? make_float4(n.x / l, n.y / l, n.z / l, n.w) : n;
            break;
        }
        case 5: v = A.cloudOnly[p]; break;
        case 6: {
            const float4 d = A.depth[p];
            v = make_float4(d.x, sqrtf(fmaxf(0.0f, d.y - d.x * d.x)), 0.0f, 1.0f);
            break;
        }
        case 7: v = make_float4(A.depth[p].z, 0.0f, 0.0f, 1.0f); break;
        case 8: v = make_float4(A.flow[p].x, A.flow[p].y, 0.0f, 1.0f); break;
        case 9: v = make_float4(A.flowRev[p].x, A.flowRev[p].y, 0.0f, 1.0f); break;
        case 10: v = make_float4(A.dens[p].z, A.dens[p].w, 0.0f, 1.0f); break;
        case 11: v = make_float4(A.depth[p].w, 0.0f, 0.0f, 1.0f); break;
        case 12: {
            const float4 d = A.dens[p];
            v = make_float4(d.x, sqrtf(fmaxf(0.0f, d.y - d.x * d.x)), 0.0f, 1.0f);
            break;
        }
        case 13: v = A.background[p]; break;
        case 14: v = make_float4(A.flow[p].z, A.flow[p].w, 0.0f, 1.0f); break;
        case 17: v = A.albedo[p]; break;
        case kFeatureMapRadianceVariance: {
            // E[x^2] - E[x]^2 of the accumulated samples (population variance of the running mean's samples)
            const float4 q = A.radianceSq[p], m = result[p];
            v = make_float4(fmaxf(0.0f, q.x - m.x * m.x), fmaxf(0.0f, q.y - m.y * m.y), fmaxf(0.0f, q.z - m.z * m.z),
                            1.0f);
            break;
        }
        default: v = make_float4(result[p].x, result[p].y, result[p].z, 1.0f); break;
    }
    outCHW[p] = v.x;
    outCHW[npix + p] = v.y;
    outCHW[2 * size_t(npix) + p] = v.z;
    if (channels == 4) outCHW[3 * size_t(npix) + p] = v.w;
}

__global__ void k_zeroU32(uint32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ------------------------------------------------------------------------------------------------------------
// Batched transmittance queries.
// ------------------------------------------------------------------------------------------------------------
template <class Grid>
CRVPT_DEV float transmittanceWalker(const SceneParams& P, const SupervoxelGrid& S, const Grid& grid, Rng& rng, V3 o,
                                    V3 d, float limit, bool binary) {
    Walker W;
    if (!walkerStart<VPT_DELTA_TRACKING>(W, P, S, o, d, limit)) return 1.0f;
    typename Grid::Cache gc;
    float T = 1.0f;
    while (true) {
        const int ev = walkStep(W, rng, P, S);
        if (ev == WALK_END) break;
        if (ev != WALK_COLLIDE) continue;
        const float rho = density(grid, P, rng, W.o + W.w * W.t, gc);
        const float pReal = P.extinction * rho / W.mu;
        if (binary) {
            if (rng.uniform() < pReal) return 0.0f;
        } else {
            T *= 1.0f - pReal;
        }
    }
    return T;
}

template <class Grid>
CRVPT_DEV float transmittanceRrt(const SceneParams& P, const SupervoxelGrid& S, const Grid& grid, Rng& rng, V3 x,
                                 V3 w, float limit) {
    const V3 eps = v3(1e-6f);
    float tMinVal, tMaxVal;
    if (!rayBoxIntersect(v3(P.boxMin) + eps, v3(P.boxMax) - eps, x, w, tMinVal, tMaxVal)) return 1.0f;
    x += w * tMinVal;
    const V3 oldX = x;
    const float lim = limit - tMinVal;
    RefDda q;
    if (!refDdaInit(P, S, x, w, tMaxVal - tMinVal, q)) return 1.0f;
    float T = 1.0f, wSum = 0.0f, resT = 0.0f, resDist = 0.0f;
    WorkCount wc;
    typename Grid::Cache gc;
    while (refInside(S, q.idx[0], q.idx[1], q.idx[2])) {
        const size_t si = refIndex(S, q.idx[0], q.idx[1], q.idx[2]);
        const float muC = S.rrt[2 * si], muR = S.rrt[2 * si + 1];
        V3 mn, mx;
        refSvBounds(P, S, q.idx, mn, mx);
        float a = 0.0f, b = 0.0f;
        rayBoxIntersect(mn, mx, oldX, w, a, b);
        b = fminf(b, lim);
        if (b > a) T *= rrtEstimator(P, grid, rng, oldX + w * a, w, a, b, T, wSum, resT, resDist, muC, muR, wc, gc);
        refDdaAdvance(q);
    }
    return T;
}

template <class Grid>
CRVPT_DEV float transmittanceDecomposition(const SceneParams& P, const SupervoxelGrid& S, const Grid& grid, Rng& rng,
                                           V3 x, V3 w, float limit) {
    const V3 eps = v3(1e-6f);
    float tMinVal, tMaxVal;
    if (!rayBoxIntersect(v3(P.boxMin) + eps, v3(P.boxMax) - eps, x, w, tMinVal, tMaxVal)) return 1.0f;
    const V3 dims = gridDims(P);
    const V3 boxDelta = v3(P.boxMax) - v3(P.boxMin);
    const V3 svSize = v3(float(S.refSize)) * boxDelta / dims;
    x += w * tMinVal;
    float travelled = tMinVal;
    typename Grid::Cache gc;
    const V3 sp = (x - v3(P.boxMin)) / boxDelta * dims / v3(float(S.refSize));
    int ix = int(floorf(sp.x)), iy = int(floorf(sp.y)), iz = int(floorf(sp.z));
    while (refInside(S, ix, iy, iz)) {
        const V3 minP = v3(P.boxMin) + v3(float(ix), float(iy), float(iz)) * svSize;
        const V3 maxP = minP + svSize;
        float a = 0.0f, b = 0.0f;
        rayBoxIntersect(minP, maxP, x, w, a, b);
        const float dMax = b - a;
        x += w * a;
        travelled += a;
        const float remaining = limit - travelled;
        if (remaining <= 0.0f) break;
        const float dUse = fminf(dMax, remaining);
        const size_t si = refIndex(S, ix, iy, iz);
        const float mmMin = S.decomp[2 * si], mmMax = S.decomp[2 * si + 1];
        if (mmMax >= 1e-5f) {
            const float muC = fmaxf(0.0000000001f, P.extinction * mmMin);
            const float muR = fmaxf(0.0000000001f, P.extinction * mmMax - muC);
            const float tc = rng.exponential() / muC;
            float tr = 0.0f;
            while (true) {
                tr += rng.exponential() / muR;
                if (tc >= dUse && tr >= dUse) break;
                if (tc <= tr) return 0.0f;
                const float rho = density(grid, P, rng, x + w * tr, gc);
                if (rng.uniform() * muR < P.extinction * rho - muC) return 0.0f;
            }
        }
        if (dUse < dMax) break;
        x = x + w * dMax;
        travelled += dMax;
        const V3 cc = (minP + maxP) * 0.5f;
        V3 mov = x + w * 0.00001f - cc;
        const V3 sm = v3(signf(mov.x), signf(mov.y), signf(mov.z));
        mov = mov * sm;
        ix += ((mov.x >= mov.y && mov.x >= mov.z) ? 1 : 0) * int(sm.x);
        iy += ((mov.y >= mov.x && mov.y >= mov.z) ? 1 : 0) * int(sm.y);
        iz += ((mov.z >= mov.x && mov.z >= mov.y) ? 1 : 0) * int(sm.z);
    }
    return 1.0f;
}

template <class Grid>
__global__ __launch_bounds__(256) void k_transmittance(SceneParams P, Grid grid, SupervoxelGrid S, int estimator,
                                                       const float* __restrict__ o, const float* __restrict__ d,
                                                       const float* __restrict__ tmax, int nRays, int nSamples,
                                                       uint32_t seed, float* __restrict__ mean,
                                                       float* __restrict__ m2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nRays) return;
    const V3 x = v3(o[3 * i], o[3 * i + 1], o[3 * i + 2]);
    const V3 w = v3(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
    const float lim = tmax ? tmax[i] : kInf;
    double s1 = 0.0, s2 = 0.0;
    for (int s = 0; s < nSamples; s++) {
        Rng rng;
        rng.init(uint32_t(i), uint32_t(s), seed ^ 0x5bd1e995u);
        float T;
        switch (estimator) {
            case T_RATIO: T = transmittanceWalker(P, S, grid, rng, x, w, lim, false); break;
            case T_RESIDUAL_RATIO: T = transmittanceRrt(P, S, grid, rng, x, w, lim); break;
            case T_DECOMPOSITION: T = transmittanceDecomposition(P, S, grid, rng, x, w, lim); break;
            default: T = transmittanceWalker(P, S, grid, rng, x, w, lim, true); break;
        }
        s1 += T;
        s2 += double(T) * T;
    }
    mean[i] = float(s1 / nSamples);
    if (m2) m2[i] = float(s2 / nSamples);
}

// ------------------------------------------------------------------------------------------------------------
// Supervoxel tables
// ------------------------------------------------------------------------------------------------------------
// Majorant: max normalised density over the supervoxel plus a one-voxel halo (the voxels a jittered or filtered
// lookup of a point inside the supervoxel can read), including the border value outside the grid.
template <class Grid>
__global__ void k_majorant(Grid grid, int svSize, int gx, int gy, int gz, float* __restrict__ maj) {
    const int si = blockIdx.x * blockDim.x + threadIdx.x;
    if (si >= gx * gy * gz) return;
    const int sx = si % gx, sy = (si / gx) % gy, sz = si / (gx * gy);
    float m = 0.0f;
    for (int k = sz * svSize - 1; k <= (sz + 1) * svSize; k++)
        for (int j = sy * svSize - 1; j <= (sy + 1) * svSize; j++)
            for (int i = sx * svSize - 1; i <= (sx + 1) * svSize; i++) m = fmaxf(m, grid.voxelAt(i, j, k));
    maj[si] = m;
}

// Flat index of the 8^3 blocks of a NanoVDB grid: per block in [b0, b0 + nb) the leaf offset the tree descent
// returns, or 0 and the tile value. Built once per uploaded grid; the accessor then replaces each descent by one load.
// Entries hold the leaf offset / 32; a leaf offset that is not 32-byte aligned or beyond 128 GiB cannot be encoded
// and sets *bad (the host then renders without the index).
__global__ void k_blockTable(NanoAccess grid, int bx0, int by0, int bz0, int nbx, int nby, int nbz,
                             uint32_t* __restrict__ leaf, float* __restrict__ tile, uint32_t* __restrict__ bad) {
    const size_t n = size_t(nbx) * nby * nbz;
    for (size_t b = size_t(blockIdx.x) * blockDim.x + threadIdx.x; b < n; b += size_t(gridDim.x) * blockDim.x) {
        const int bi = int(b % size_t(nbx)), bj = int((b / size_t(nbx)) % size_t(nby)),
                  bk = int(b / (size_t(nbx) * nby));
        float t = 0.0f;
        const uint64_t off = grid.locateBlock((bx0 + bi) * 8, (by0 + bj) * 8, (bz0 + bk) * 8, t);
        if ((off & 31u) != 0u || (off >> 37) != 0u) atomicOr(bad, 1u);
        leaf[b] = uint32_t(off >> 5);
        tile[b] = off ? 0.0f : t;
    }
}

// Reference tables (SuperVoxelGrid.cpp:101-211, 269-405): min, max and average of the normalised density over
// S^3 (NEAREST) or the (S+2)^3 halo (other modes), zero outside the grid, summed in the reference's z, y, x order.
template <class Grid>
__global__ void k_refSupervoxelStats(Grid grid, int dimX, int dimY, int dimZ, int svSize, int gx, int gy, int gz,
                                     int halo, float* __restrict__ mn, float* __restrict__ mx,
                                     float* __restrict__ avg) {
    const int si = blockIdx.x * blockDim.x + threadIdx.x;
    if (si >= gx * gy * gz) return;
    const int sx = si % gx, sy = (si / gx) % gy, sz = si / (gx * gy);
    float dmin = 3.402823466e38f, dmax = -3.402823466e38f, sum = 0.0f;
    int count = 0;
    const int lo = halo ? -1 : 0, hi = halo ? svSize : svSize - 1;
    for (int oz = lo; oz <= hi; oz++)
        for (int oy = lo; oy <= hi; oy++)
            for (int ox = lo; ox <= hi; ox++) {
                const int x = sx * svSize + ox, y = sy * svSize + oy, z = sz * svSize + oz;
                float v = 0.0f;
                if (x >= 0 && y >= 0 && z >= 0 && x < dimX && y < dimY && z < dimZ) v = grid.voxelAt(x, y, z);
                dmin = fminf(dmin, v);
                dmax = fmaxf(dmax, v);
                sum += v;
                count++;
            }
    mn[si] = dmin;
    mx[si] = dmax;
    avg[si] = sum / float(count);
}

// ------------------------------------------------------------------------------------------------------------
// Procedural clouds: seeded fBm value noise (the build's generator; parameters documented with the entry point).
// Contraction is off so every FP32 operation rounds exactly as in the CPU restatement of the generator.
// ------------------------------------------------------------------------------------------------------------
CRVPT_DEV uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

CRVPT_DEV float latticeValue(int ix, int iy, int iz, uint32_t salt) {
    const uint32_t h = lowbias32((uint32_t(ix) * 73856093u) ^ (uint32_t(iy) * 19349663u) ^
                                 (uint32_t(iz) * 83492791u) ^ salt);
    return float(h >> 8) * (1.0f / 16777216.0f);
}

CRVPT_DEV float fadeq(float t) {
#pragma clang fp contract(off)
    return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f);
}

CRVPT_DEV float valueNoise(float px, float py, float pz, uint32_t salt) {
#pragma clang fp contract(off)
    const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
    const int ix = int(fx), iy = int(fy), iz = int(fz);
    const float u = fadeq(px - fx), v = fadeq(py - fy), w = fadeq(pz - fz);
    float c[8];
#pragma unroll
    for (int k = 0; k < 8; k++) c[k] = latticeValue(ix + (k & 1), iy + ((k >> 1) & 1), iz + ((k >> 2) & 1), salt);
    const float x00 = c[0] + (c[1] - c[0]) * u;
    const float x10 = c[2] + (c[3] - c[2]) * u;
    const float x01 = c[4] + (c[5] - c[4]) * u;
    const float x11 = c[6] + (c[7] - c[6]) * u;
    const float y0 = x00 + (x10 - x00) * v;
    const float y1 = x01 + (x11 - x01) * v;
    return y0 + (y1 - y0) * w;
}

__global__ void k_fbm(float* __restrict__ out, int nx, int ny, int nz, FbmParams F) {
#pragma clang fp contract(off)
    const size_t total = size_t(nx) * ny * nz;
    for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
         idx += size_t(gridDim.x) * blockDim.x) {
        const int x = int(idx % nx);
        const int y = int((idx / nx) % ny);
        const int z = int(idx / (size_t(nx) * ny));
        const float px = __fdiv_rn(float(x) + 0.5f, float(nx));
        const float py = __fdiv_rn(float(y) + 0.5f, float(ny));
        const float pz = __fdiv_rn(float(z) + 0.5f, float(nz));
        float sum = 0.0f, norm = 0.0f, amp = 1.0f, freq = F.baseFrequency;
        for (int o = 0; o < F.octaves; o++) {
            const uint32_t salt = F.seed * 2654435761u + uint32_t(o) * 0x9E3779B9u;
            sum = sum + amp * valueNoise(px * freq, py * freq, pz * freq, salt);
            norm = norm + amp;
            amp = amp * F.gain;
            freq = freq * F.lacunarity;
        }
        float v = __fdiv_rn(sum, norm);
        if (F.heightFalloff) {
            // anvil profile: flat top, rounded base, in the normalised height py
            const float h = py;
            const float base = fminf(fmaxf(__fdiv_rn(h - 0.05f, 0.25f), 0.0f), 1.0f);
            const float top = fminf(fmaxf(__fdiv_rn(0.95f - h, 0.1f), 0.0f), 1.0f);
            const float spread = 1.0f + 0.5f * fminf(fmaxf(__fdiv_rn(h - 0.6f, 0.3f), 0.0f), 1.0f);
            v = v * base * top * spread;
        }
        float t = __fdiv_rn(v - F.edge0, F.edge1 - F.edge0);
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        out[idx] = t * t * (3.0f - 2.0f * t);
    }
}

// Dense field of a NanoVDB grid over its index bounding box, x fastest (CloudData::getDenseDensityField,
// CloudData.cpp:1173-1253: tree.getValue at indexBBox.min + (x, y, z)). Each lane decodes a run of consecutive x
// through one lookup memo.
constexpr int kDecodeRun = 8;
__global__ void k_decodeDense(NanoAccess grid, int nx, int ny, int nz, float* __restrict__ out) {
    const size_t runsX = (size_t(nx) + kDecodeRun - 1) / kDecodeRun;
    const size_t runs = runsX * size_t(ny) * nz;
    for (size_t r = size_t(blockIdx.x) * blockDim.x + threadIdx.x; r < runs; r += size_t(gridDim.x) * blockDim.x) {
        const int x0 = int(r % runsX) * kDecodeRun;
        const size_t yz = r / runsX;
        const int y = int(yz % size_t(ny)), z = int(yz / size_t(ny));
        LeafCache c;
        const int x1 = min(x0 + kDecodeRun, nx);
        for (int x = x0; x < x1; x++) out[(size_t(z) * ny + y) * nx + x] = grid.voxelAt(x, y, z, c);
    }
}

// Dense grid normalisation: (raw - vmin) / (vmax - vmin) per voxel, as sampleGridImage applies it per lookup.
__global__ void k_normalize(const float* __restrict__ raw, float* __restrict__ out, size_t n, float vmin,
                            float vmax) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        out[i] = __fdiv_rn(raw[i] - vmin, vmax - vmin);
}

// ------------------------------------------------------------------------------------------------------------
// Grid lookup probes (exactness checks of the accessors)
// ------------------------------------------------------------------------------------------------------------
constexpr int kLookupRun = 8;

template <class Grid>
__global__ void k_gridLookup(Grid grid, int n, const int32_t* __restrict__ ijk, float* __restrict__ out) {
    // Each lane answers a run of kLookupRun consecutive queries through one lookup memo, the way a path's
    // consecutive density taps do, so coherent query lists exercise memo hits as well as misses.
    const int first = (blockIdx.x * blockDim.x + threadIdx.x) * kLookupRun;
    const int last = min(first + kLookupRun, n);
    typename Grid::Cache c;
    for (int q = first; q < last; q++) out[q] = grid.voxelAt(ijk[3 * q], ijk[3 * q + 1], ijk[3 * q + 2], c);
}

// NanoVDB world (grid) coordinates -> world_to_indexf -> floor -> value, the nearest sampleCloudRaw path.
__global__ void k_nanoLookupWorld(NanoAccess grid, int n, const float* __restrict__ xyz, float* __restrict__ out,
                                  int32_t* __restrict__ ijkOut) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V3 idx = grid.worldToIndex(v3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
    const int a = int(floorf(idx.x)), b = int(floorf(idx.y)), c = int(floorf(idx.z));
    ijkOut[3 * i] = a;
    ijkOut[3 * i + 1] = b;
    ijkOut[3 * i + 2] = c;
    out[i] = grid.value(a, b, c);
}

// ------------------------------------------------------------------------------------------------------------
// Host-side launchers
// ------------------------------------------------------------------------------------------------------------
template <int MODE, class Grid>
static void launchPool(dim3 grid, dim3 block, hipStream_t stream, const SceneParams& P, const Grid& g,
                       const SupervoxelGrid& S, float4* paths, uint32_t* counter, int steps,
                       unsigned long long* stats, float4* feat) {
    if (stats)
        hipLaunchKernelGGL((k_pathPool<MODE, Grid, true>), grid, block, 0, stream, P, g, S, paths, counter, steps,
                           stats, feat);
    else
        hipLaunchKernelGGL((k_pathPool<MODE, Grid, false>), grid, block, 0, stream, P, g, S, paths, counter, steps,
                           stats, feat);
}

template <class Grid>
static hipError_t launchRenderT(const SceneParams& P, const Grid& grid, const SupervoxelGrid& S, float4* paths,
                                uint32_t* counter, const LaunchConfig& cfg, unsigned long long* stats, float4* feat,
                                hipStream_t stream) {
    hipLaunchKernelGGL(k_zeroU32, dim3((kClaimParts * kCounterStride + 255) / 256), dim3(256), 0, stream, counter,
                       int(kClaimParts * kCounterStride));
    const dim3 block(256);
    const dim3 gridDim(cfg.numBlocks);
    const int st = cfg.stepsPerRefill;
    switch (P.mode) {
        case VPT_RATIO_TRACKING:
            launchPool<VPT_RATIO_TRACKING>(gridDim, block, stream, P, grid, S, paths, counter, st, stats, feat);
            break;
        case VPT_NEXT_EVENT_TRACKING:
            launchPool<VPT_NEXT_EVENT_TRACKING>(gridDim, block, stream, P, grid, S, paths, counter, st, stats, feat);
            break;
        case VPT_SPECTRAL_DELTA_TRACKING:
            launchPool<VPT_SPECTRAL_DELTA_TRACKING>(gridDim, block, stream, P, grid, S, paths, counter, st, stats,
                                                    feat);
            break;
        case VPT_NEXT_EVENT_TRACKING_SPECTRAL:
            launchPool<VPT_NEXT_EVENT_TRACKING_SPECTRAL>(gridDim, block, stream, P, grid, S, paths, counter, st,
                                                         stats, feat);
            break;
        case VPT_DECOMPOSITION_TRACKING:
            launchPool<VPT_DECOMPOSITION_TRACKING>(gridDim, block, stream, P, grid, S, paths, counter, st, stats,
                                                   feat);
            break;
        case VPT_RESIDUAL_RATIO_TRACKING:
            launchPool<VPT_RESIDUAL_RATIO_TRACKING>(gridDim, block, stream, P, grid, S, paths, counter, st, stats,
                                                    feat);
            break;
        default:
            launchPool<VPT_DELTA_TRACKING>(gridDim, block, stream, P, grid, S, paths, counter, st, stats, feat);
            break;
    }
    return hipGetLastError();
}

__global__ void k_storeParams(SceneParams* __restrict__ dst, SceneParams P) {
    // word by word from the kernel argument (the struct is a whole number of 4-byte words)
    static_assert(sizeof(SceneParams) % 4 == 0, "SceneParams is copied in 4-byte words");
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&P);
    uint32_t* out = reinterpret_cast<uint32_t*>(dst);
    for (uint32_t i = threadIdx.x; i < sizeof(SceneParams) / 4; i += blockDim.x) out[i] = src[i];
}

hipError_t launchStoreParams(SceneParams* dst, const SceneParams& P, hipStream_t stream) {
    hipLaunchKernelGGL(k_storeParams, dim3(1), dim3(256), 0, stream, dst, P);
    return hipGetLastError();
}

uint32_t maxFramesPerLaunch(int width, int height) {
    // every claim index (screen tiles x tile pixels x frames, padding included) stays below 2^31 (udiv31)
    const uint64_t tile = kClaimTile == 0u ? 1u : kClaimTile;
    const uint64_t perFrame = ((uint64_t(width) + tile - 1u) / tile) * ((uint64_t(height) + tile - 1u) / tile) *
                              tile * tile;
    return uint32_t(std::min<uint64_t>(0x7FFFFFFFu / std::max<uint64_t>(perFrame, 1u), 0xFFFFFFFFu));
}

hipError_t launchRender(const SceneParams& Pin, const DenseGrid* dense, const NanoGrid* nano, const SupervoxelGrid& S,
                        float4* paths, uint32_t* counter, const LaunchConfig& cfg, unsigned long long* stats,
                        float4* feat, hipStream_t stream) {
    if (Pin.numFrames == 0u || Pin.numFrames > maxFramesPerLaunch(Pin.width, Pin.height) || !Pin.mem)
        return hipErrorInvalidValue;
    SceneParams P = Pin;
    const uint32_t tp = kClaimTile == 0u ? 1u : kClaimTile * kClaimTile;
    P.divTileClaims = makeDiv31(tp * P.numFrames);
    P.divTilesX = makeDiv31(kClaimTile == 0u ? 1u : (uint32_t(P.width) + kClaimTile - 1u) / kClaimTile);
    // this launch's constants, with its frame range and claim divisors, in the device copy the kernels read
    if (const hipError_t e = launchStoreParams(const_cast<SceneParams*>(Pin.mem), P, stream)) return e;
    // variants: 64-bit offsets and the full index map only for NanoVDB grids of 4 GiB and more or with rotated
    // voxels, 64-bit voxel indices only for dense grids of 2^32 voxels and more; the trilinear code only for the
    // trilinear filter
    const bool tri = P.interpolation == INTERP_TRILINEAR;
    if (nano && (nano->wide || !nano->tapDiag))
        return launchRenderT(P, NanoAccess{*nano}, S, paths, counter, cfg, stats, feat, stream);
    if (nano && tri) return launchRenderT(P, NanoAccessNarrow{*nano}, S, paths, counter, cfg, stats, feat, stream);
    if (nano) return launchRenderT(P, NanoAccessNarrowPoint{*nano}, S, paths, counter, cfg, stats, feat, stream);
    const bool small =
            uint64_t(dense->dim[0]) * uint64_t(dense->dim[1]) * uint64_t(dense->dim[2]) < (uint64_t(1) << 32);
    if (!tri && small) return launchRenderT(P, DenseAccessPoint{*dense}, S, paths, counter, cfg, stats, feat, stream);
    return launchRenderT(P, DenseAccess{*dense}, S, paths, counter, cfg, stats, feat, stream);
}

hipError_t launchAccumulate(const float4* paths, uint32_t npix, uint32_t frames, float4* acc, uint32_t prevFrames,
                            float* outCHW, int channels, int foreground, hipStream_t stream) {
    hipLaunchKernelGGL(k_accumulate, dim3((npix + 255) / 256), dim3(256), 0, stream, paths, npix, frames, acc,
                       prevFrames, outCHW, channels, foreground);
    return hipGetLastError();
}

hipError_t launchAccumulateFeatures(const SceneParams& P, const float4* paths, const float4* feat, uint32_t npix,
                                    uint32_t frames, uint32_t prevFrames, bool noAccumulation, const FeatureAcc& acc,
                                    hipStream_t stream) {
    hipLaunchKernelGGL(k_accumulateFeatures, dim3((npix + 255) / 256), dim3(256), 0, stream, P, paths, feat, npix,
                       frames, prevFrames, int(noAccumulation), acc);
    return hipGetLastError();
}

hipError_t launchFeatureOut(int map, const FeatureAcc& acc, const float4* result, uint32_t npix, float* outCHW,
                            int channels, hipStream_t stream) {
    hipLaunchKernelGGL(k_featureOut, dim3((npix + 255) / 256), dim3(256), 0, stream, map, acc, result, npix, outCHW,
                       channels);
    return hipGetLastError();
}

hipError_t launchWriteCHW(const float4* img, uint32_t npix, float* outCHW, int channels, hipStream_t stream) {
    hipLaunchKernelGGL(k_writeCHW, dim3((npix + 255) / 256), dim3(256), 0, stream, img, npix, outCHW, channels);
    return hipGetLastError();
}

hipError_t launchTransmittance(const SceneParams& P, const DenseGrid* dense, const NanoGrid* nano,
                               const SupervoxelGrid& S, int estimator, const float* o, const float* d,
                               const float* tmax, int nRays, int nSamples, uint32_t seed, float* mean, float* m2,
                               hipStream_t stream) {
    const dim3 grid((nRays + 255) / 256), block(256);
    if (nano) {
        hipLaunchKernelGGL(k_transmittance<NanoAccess>, grid, block, 0, stream, P, NanoAccess{*nano}, S, estimator,
                           o, d, tmax, nRays, nSamples, seed, mean, m2);
    } else {
        hipLaunchKernelGGL(k_transmittance<DenseAccess>, grid, block, 0, stream, P, DenseAccess{*dense}, S,
                           estimator, o, d, tmax, nRays, nSamples, seed, mean, m2);
    }
    return hipGetLastError();
}

hipError_t launchBlockTable(const NanoGrid& nano, const int b0[3], const int nb[3], uint32_t* leaf, float* tile,
                            uint32_t* bad, hipStream_t stream) {
    NanoGrid g = nano;
    g.blockLeaf = nullptr;  // the table is built from the tree itself
    g.blockTile = nullptr;
    const size_t n = size_t(nb[0]) * nb[1] * nb[2];
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_blockTable, dim3(unsigned(blocks)), dim3(256), 0, stream, NanoAccess{g}, b0[0], b0[1], b0[2],
                       nb[0], nb[1], nb[2], leaf, tile, bad);
    return hipGetLastError();
}

hipError_t launchMajorant(const DenseGrid* dense, const NanoGrid* nano, const int dims[3], int svSize,
                          const int grid3[3], float* maj, hipStream_t stream) {
    (void)dims;
    const int n = grid3[0] * grid3[1] * grid3[2];
    if (nano) {
        hipLaunchKernelGGL(k_majorant<NanoAccess>, dim3((n + 127) / 128), dim3(128), 0, stream, NanoAccess{*nano},
                           svSize, grid3[0], grid3[1], grid3[2], maj);
    } else {
        hipLaunchKernelGGL(k_majorant<DenseAccess>, dim3((n + 127) / 128), dim3(128), 0, stream,
                           DenseAccess{*dense}, svSize, grid3[0], grid3[1], grid3[2], maj);
    }
    return hipGetLastError();
}

hipError_t launchRefSupervoxelStats(const DenseGrid* dense, const NanoGrid* nano, const int dims[3], int svSize,
                                    const int grid3[3], int halo, float* mn, float* mx, float* avg,
                                    hipStream_t stream) {
    const int n = grid3[0] * grid3[1] * grid3[2];
    if (nano) {
        hipLaunchKernelGGL(k_refSupervoxelStats<NanoAccess>, dim3((n + 63) / 64), dim3(64), 0, stream,
                           NanoAccess{*nano}, dims[0], dims[1], dims[2], svSize, grid3[0], grid3[1], grid3[2], halo,
                           mn, mx, avg);
    } else {
        hipLaunchKernelGGL(k_refSupervoxelStats<DenseAccess>, dim3((n + 63) / 64), dim3(64), 0, stream,
                           DenseAccess{*dense}, dims[0], dims[1], dims[2], svSize, grid3[0], grid3[1], grid3[2],
                           halo, mn, mx, avg);
    }
    return hipGetLastError();
}

hipError_t launchFbm(float* out, int nx, int ny, int nz, const FbmParams& F, hipStream_t stream) {
    const size_t total = size_t(nx) * ny * nz;
    const size_t blocks = std::min<size_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fbm, dim3(unsigned(blocks)), dim3(256), 0, stream, out, nx, ny, nz, F);
    return hipGetLastError();
}

hipError_t launchDecodeDense(const NanoGrid& nano, const int dims[3], float* out, hipStream_t stream) {
    const size_t runs = (size_t(dims[0]) + kDecodeRun - 1) / kDecodeRun * size_t(dims[1]) * size_t(dims[2]);
    const size_t blocks = std::min<size_t>((runs + 255) / 256, 65536);
    hipLaunchKernelGGL(k_decodeDense, dim3(unsigned(blocks)), dim3(256), 0, stream, NanoAccess{nano}, dims[0], dims[1],
                       dims[2], out);
    return hipGetLastError();
}

hipError_t launchNormalize(const float* raw, float* out, size_t n, float vmin, float vmax, hipStream_t stream) {
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_normalize, dim3(unsigned(blocks)), dim3(256), 0, stream, raw, out, n, vmin, vmax);
    return hipGetLastError();
}

hipError_t launchGridLookup(const DenseGrid* dense, const NanoGrid* nano, int n, const int32_t* ijk, float* out,
                            hipStream_t stream) {
    const dim3 grid((n + 256 * kLookupRun - 1) / (256 * kLookupRun)), block(256);
    if (nano) {
        // index-space coordinates: voxelAt adds the index minimum, so pass raw NanoVDB coordinates through value()
        NanoGrid g = *nano;
        g.indexMin[0] = g.indexMin[1] = g.indexMin[2] = 0;
        hipLaunchKernelGGL(k_gridLookup<NanoAccess>, grid, block, 0, stream, NanoAccess{g}, n, ijk, out);
    } else {
        hipLaunchKernelGGL(k_gridLookup<DenseAccess>, grid, block, 0, stream, DenseAccess{*dense}, n, ijk, out);
    }
    return hipGetLastError();
}

hipError_t launchNanoLookupWorld(const NanoGrid& nano, int n, const float* xyz, float* out, int32_t* ijk,
                                 hipStream_t stream) {
    hipLaunchKernelGGL(k_nanoLookupWorld, dim3((n + 255) / 256), dim3(256), 0, stream, NanoAccess{nano}, n, xyz, out,
                       ijk);
    return hipGetLastError();
}

template <class Grid>
static const void* pathPoolKernel(int mode) {
    switch (mode) {
        case VPT_RATIO_TRACKING: return (const void*)k_pathPool<VPT_RATIO_TRACKING, Grid, false>;
        case VPT_NEXT_EVENT_TRACKING: return (const void*)k_pathPool<VPT_NEXT_EVENT_TRACKING, Grid, false>;
        case VPT_SPECTRAL_DELTA_TRACKING: return (const void*)k_pathPool<VPT_SPECTRAL_DELTA_TRACKING, Grid, false>;
        case VPT_NEXT_EVENT_TRACKING_SPECTRAL:
            return (const void*)k_pathPool<VPT_NEXT_EVENT_TRACKING_SPECTRAL, Grid, false>;
        case VPT_DECOMPOSITION_TRACKING: return (const void*)k_pathPool<VPT_DECOMPOSITION_TRACKING, Grid, false>;
        case VPT_RESIDUAL_RATIO_TRACKING: return (const void*)k_pathPool<VPT_RESIDUAL_RATIO_TRACKING, Grid, false>;
        default: return (const void*)k_pathPool<VPT_DELTA_TRACKING, Grid, false>;
    }
}

int occupancyBlocksPerCU(int mode, bool sparse) {
    int blocks = 0;
    const void* k = sparse ? pathPoolKernel<NanoAccessNarrowPoint>(mode) : pathPoolKernel<DenseAccessPoint>(mode);
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 256, 0);
    if (e != hipSuccess || blocks <= 0) return 4;
    return blocks;
}

}  // namespace crvpt
