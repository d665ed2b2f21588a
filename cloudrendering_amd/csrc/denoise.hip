// Edge-avoiding a-trous wavelet denoiser (EAWDenoise.glsl:161-314 compute path, EAWDenoiser.cpp:345-391 host
// loop): maxNumIterations 5x5 passes with step widths 1, 2, 4, ..., Gaussian kernel exp(-(x^2 + y^2) / 2) and
// edge-stopping weights from colour, position and normal. The reference compute path uses the constant colour
// weight exp(-1 / phiColor) (independent of the colour distance); useColorDistance = 1 selects the paper's
// exp(-|dc|^2 / phiColor) instead. Taps outside the image are clamped to the edge.
#include <hip/hip_runtime.h>

#include "render_api.hpp"

namespace crvpt {

// One pass with step width s on a dilated lattice: a workgroup takes the 16 x 16 pixels (px + s i, py + s j) of one
// phase (px, py) = (x mod s, y mod s) of the image, whose 5 x 5 taps are the lattice points of a 20 x 20 tile of the
// same phase. The tile (colour, position, normal) is staged in LDS once per workgroup, 400 loads per 256 pixels,
// instead of 25 global taps per pixel; lattice points outside the image hold the clamped edge pixel, which is what
// the reference's clamped fetch reads. The weights are evaluated per tap in the same operations and order as a
// direct 25-tap loop.
constexpr int kEawTile = 16;
constexpr int kEawHalo = kEawTile + 4;

__global__ __launch_bounds__(256) void k_eawPass(const float4* __restrict__ color, const float4* __restrict__ pos,
                                                 const float4* __restrict__ nrm, float4* __restrict__ out, int w,
                                                 int h, int stepWidth, float phiColor, float phiPosition,
                                                 float phiNormal, int useColorDistance) {
    __shared__ float4 tc[kEawHalo][kEawHalo];
    __shared__ float4 tp[kEawHalo][kEawHalo];
    __shared__ float4 tn[kEawHalo][kEawHalo];
    const int s = stepWidth;
    // blockIdx.x = tile column * s + phase x, blockIdx.y = tile row * s + phase y
    const int phx = int(blockIdx.x) % s, phy = int(blockIdx.y) % s;
    const int lx0 = (int(blockIdx.x) / s) * kEawTile - 2, ly0 = (int(blockIdx.y) / s) * kEawTile - 2;
    const int tid = threadIdx.y * kEawTile + threadIdx.x;
    for (int e = tid; e < kEawHalo * kEawHalo; e += kEawTile * kEawTile) {
        const int i = e % kEawHalo, j = e / kEawHalo;
        const int gx = min(max(phx + (lx0 + i) * s, 0), w - 1);
        const int gy = min(max(phy + (ly0 + j) * s, 0), h - 1);
        const size_t g = size_t(gy) * w + gx;
        tc[j][i] = color[g];
        if (pos) tp[j][i] = pos[g];
        if (nrm) tn[j][i] = nrm[g];
    }
    __syncthreads();
    const int x = phx + (lx0 + 2 + int(threadIdx.x)) * s;
    const int y = phy + (ly0 + 2 + int(threadIdx.y)) * s;
    if (x >= w || y >= h) return;
    const int cx = threadIdx.x + 2, cy = threadIdx.y + 2;
    const float4 c0 = tc[cy][cx];
    const float4 p0 = pos ? tp[cy][cx] : make_float4(0, 0, 0, 0);
    const float4 n0 = nrm ? tn[cy][cx] : make_float4(0, 0, 0, 0);
    const float constColorWeight = fminf(__expf(-1.0f / phiColor), 1.0f);
    // exp(-|d|^2 / phi) as exp(|d|^2 * (-1 / phi)): one multiply per weight instead of an IEEE division
    const float nic = -1.0f / phiColor, nip = -1.0f / phiPosition, nin = -1.0f / phiNormal;
    float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float accW = 0.0f;
#pragma unroll
    for (int i = 0; i < 25; i++) {
        const int ox = i % 5 - 2, oy = i / 5 - 2;
        const float k = __expf(-float(ox * ox + oy * oy) / 2.0f);
        const float4 c = tc[cy + oy][cx + ox];
        float wgt;
        if (useColorDistance) {
            const float dx = c0.x - c.x, dy = c0.y - c.y, dz = c0.z - c.z, dw = c0.w - c.w;
            wgt = fminf(__expf((dx * dx + dy * dy + dz * dz + dw * dw) * nic), 1.0f);
        } else {
            wgt = constColorWeight;
        }
        if (pos) {
            const float4 p = tp[cy + oy][cx + ox];
            const float dx = p0.x - p.x, dy = p0.y - p.y, dz = p0.z - p.z, dw = p0.w - p.w;
            wgt *= fminf(__expf((dx * dx + dy * dy + dz * dz + dw * dw) * nip), 1.0f);
        }
        if (nrm) {
            const float4 n = tn[cy + oy][cx + ox];
            const float dx = n0.x - n.x, dy = n0.y - n.y, dz = n0.z - n.z, dw = n0.w - n.w;
            wgt *= fminf(__expf((dx * dx + dy * dy + dz * dz + dw * dw) * nin), 1.0f);
        }
        sum.x += c.x * wgt * k;
        sum.y += c.y * wgt * k;
        sum.z += c.z * wgt * k;
        sum.w += c.w * wgt * k;
        accW += wgt * k;
    }
    out[size_t(y) * w + x] = make_float4(sum.x / accW, sum.y / accW, sum.z / accW, sum.w / accW);
}

hipError_t launchEaw(const float4* color, const float4* position, const float4* normal, float4* ping, float4* pong,
                     const EawParams& p, float4** result, hipStream_t stream) {
    const dim3 block(kEawTile, kEawTile);
    const float4* src = color;
    float4* bufs[2] = {ping, pong};
    int stepWidth = 1;
    for (int i = 0; i < p.iterations; i++) {
        float4* dst = bufs[i & 1];
        // lattice points per phase: ceil(w / s) x ceil(h / s), in 16 x 16 tiles, for each of the s x s phases
        const int tilesX = ((p.width + stepWidth - 1) / stepWidth + kEawTile - 1) / kEawTile;
        const int tilesY = ((p.height + stepWidth - 1) / stepWidth + kEawTile - 1) / kEawTile;
        const dim3 grid(unsigned(tilesX * stepWidth), unsigned(tilesY * stepWidth));
        hipLaunchKernelGGL(k_eawPass, grid, block, 0, stream, src, position, normal, dst, p.width, p.height, stepWidth,
                           p.phiColor, p.phiPosition, p.phiNormal, p.useColorDistance);
        src = dst;
        stepWidth *= 2;
    }
    *result = const_cast<float4*>(src);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------------------
// SVGF (Data/Shaders/Denoiser/SVGF/SVGF.glsl, svgf_common.glsl; sequence of SVGFDenoiser::denoise,
// src/Denoiser/SVGF.cpp:118-180). Inputs are the renderer's non-accumulated maps: noisy colour, normal, depth,
// depth fwidth and motion (the Flow map). Out-of-image fetches read 0 (texelFetch outside the texture).
// ------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ bool inImage(int x, int y, int w, int h) { return x >= 0 && y >= 0 && x < w && y < h; }
__device__ __forceinline__ float4 fetch4(const float4* img, int x, int y, int w, int h) {
    return inImage(x, y, w, h) ? img[size_t(y) * w + x] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
__device__ __forceinline__ float fetchX(const float4* img, int x, int y, int w, int h) {
    return inImage(x, y, w, h) ? img[size_t(y) * w + x].x : 0.0f;
}
__device__ __forceinline__ float fetchW(const float4* img, int x, int y, int w, int h) {
    return inImage(x, y, w, h) ? img[size_t(y) * w + x].w : 0.0f;
}

// compute_weight (svgf_common.glsl): sigma_n = 128; fmaxf keeps the non-NaN operand like GLSL max on the GPU.
__device__ __forceinline__ float svgfWeight(float cd, float od, float phiD, float4 cn, float4 on, float cc, float oc,
                                            float phiC) {
    const float d = fmaxf(0.0f, cn.x * on.x + cn.y * on.y + cn.z * on.z);
    const float wn = __powf(d, 128.0f);
    const float wz = phiD == 0.0f ? 0.0f : fabsf(cd - od) / phiD;
    const float wc = fabsf(cc - oc) * 2.0f / phiC;
    return __expf(0.0f - fmaxf(wc, 0.0f) - fmaxf(wz, 0.0f)) * wn;
}

struct SvgfFrame {
    const float4* noisy;    // colour of this frame (rgb)
    const float4* flow;     // motion in .xy (pixels)
    const float4* normal;   // .xyz
    const float4* depth;    // depth in .x, depth fwidth in .w
    const float4* prevNormal;
    const float4* prevDepth;  // .x
    const float4* colorHistory;
    const float4* momentsHistory;  // (moments, history length, 0)
    int w, h;
    float allowedZ, allowedN;
};

__device__ __forceinline__ bool svgfValid(const SvgfFrame& F, int x, int y, float z, float4 n) {
    if (x < 1 || y < 1 || x > F.w - 1 || y > F.h - 1) return false;
    const float zp = fetchX(F.prevDepth, x, y, F.w, F.h);
    const float4 np = fetch4(F.prevNormal, x, y, F.w, F.h);
    if (fabsf(zp - z) > F.allowedZ) return false;
    const float dx = np.x - n.x, dy = np.y - n.y, dz = np.z - n.z;
    if (sqrtf(dx * dx + dy * dy + dz * dz) > F.allowedN) return false;
    return true;
}

// Compute-Reproject: temporal accumulation of colour and first two luminance (red channel) moments.
__global__ void k_svgfReproject(SvgfFrame F, float4* __restrict__ accumMoments, float4* __restrict__ tempAccum) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= F.w || y >= F.h) return;
    const size_t i = size_t(y) * F.w + x;
    const float4 mo = F.flow[i];
    const int px = int(0.5f + float(x) - mo.x), py = int(0.5f + float(y) - mo.y);
    bool success = inImage(px, py, F.w, F.h);
    float2 prevM = make_float2(0.0f, 0.0f);
    float histLen = 0.0f;
    if (success) {
        const float4 m = F.momentsHistory[size_t(py) * F.w + px];
        prevM = make_float2(m.x, m.y);
        histLen = m.z;
    }
    const float4 c4 = F.noisy[i];
    float4 clf = F.colorHistory[i];
    if (success) {
        const float fx = 0.01f + float(x) - mo.x, fy = 0.01f + float(y) - mo.y;
        const int ix = int(fx), iy = int(fy);
        const float z = F.depth[i].x;
        const float4 n = F.normal[i];
        // try_2x2_tap: offsets (0,0), (0,1), (1,0), (1,1) paired with the weights of (0,0), (1,0), (0,1), (1,1)
        const int ox[4] = {0, 0, 1, 1}, oy[4] = {0, 1, 0, 1};
        const float ax = fx - floorf(fx), ay = fy - floorf(fy);
        const float wts[4] = {(1 - ax) * (1 - ay), ax * (1 - ay), (1 - ax) * ay, ax * ay};
        float3 cb = make_float3(0.0f, 0.0f, 0.0f);
        float2 mb = make_float2(0.0f, 0.0f);
        float sw = 0.0f;
        bool any = false;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!svgfValid(F, ix + ox[k], iy + oy[k], z, n)) continue;
            any = true;
            const float4 oc = fetch4(F.colorHistory, ix + ox[k], iy + oy[k], F.w, F.h);
            const float4 om = fetch4(F.momentsHistory, ix + ox[k], iy + oy[k], F.w, F.h);
            cb = make_float3(cb.x + wts[k] * oc.x, cb.y + wts[k] * oc.y, cb.z + wts[k] * oc.z);
            mb = make_float2(mb.x + wts[k] * om.x, mb.y + wts[k] * om.y);
            sw += wts[k];
        }
        success = any && sw >= 0.001f;
        if (success) {
            clf = make_float4(cb.x / sw, cb.y / sw, cb.z / sw, 0.0f);
            prevM = make_float2(mb.x / sw, mb.y / sw);
        } else {
            // try_3x3_bilat
            float3 fc = make_float3(0.0f, 0.0f, 0.0f);
            float2 fm = make_float2(0.0f, 0.0f);
            float nv = 0.0f;
            for (int yy = -1; yy <= 1; yy++)
                for (int xx = -1; xx <= 1; xx++) {
                    const int qx = ix + xx, qy = iy + yy;
                    if (qx < 1 || qy < 1 || qx >= F.w || qy >= F.h) continue;
                    if (!svgfValid(F, qx, qy, z, n)) continue;
                    const float4 oc = F.colorHistory[size_t(qy) * F.w + qx];
                    const float4 om = F.momentsHistory[size_t(qy) * F.w + qx];
                    fc = make_float3(fc.x + oc.x, fc.y + oc.y, fc.z + oc.z);
                    fm = make_float2(fm.x + om.x, fm.y + om.y);
                    nv += 1.0f;
                }
            if (nv > 0.0f) {
                success = true;
                clf = make_float4(fc.x / nv, fc.y / nv, fc.z / nv, 0.0f);
                prevM = make_float2(fm.x / nv, fm.y / nv);
            }
        }
    }
    const float hl = fminf(success ? histLen + 1.0f : 1.0f, 32.0f);
    const float ac = success ? fmaxf(0.01f, 1.0f / hl) : 1.0f;
    const float am = success ? fmaxf(0.2f, 1.0f / hl) : 1.0f;
    const float m0 = c4.x, m1 = m0 * m0;
    const float mr = prevM.x * (1.0f - am) + m0 * am;
    const float mg = prevM.y * (1.0f - am) + m1 * am;
    const float var = fmaxf(0.0f, mg - mr * mr);
    accumMoments[i] = make_float4(mr, mg, hl, 0.0f);
    tempAccum[i] = make_float4(clf.x * (1.0f - ac) + c4.x * ac, clf.y * (1.0f - ac) + c4.y * ac,
                               clf.z * (1.0f - ac) + c4.z * ac, var);
}

// Compute-Filter-Moments: 7x7 cross-bilateral estimate of the variance while the history is short (< 4 frames).
__global__ void k_svgfFilterMoments(SvgfFrame F, const float4* __restrict__ accumMoments,
                                    const float4* __restrict__ tempIn, float4* __restrict__ tempOut) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= F.w || y >= F.h) return;
    const size_t i = size_t(y) * F.w + x;
    const float hl = accumMoments[i].z;
    const float4 cc = tempIn[i];
    if (hl >= 4.0f) {
        tempOut[i] = cc;
        return;
    }
    const float cd = F.depth[i].x;
    const float cf = fabsf(F.depth[i].w) + 0.0001f;
    const float4 cn = F.normal[i];
    float sw = 0.0f;
    float3 sc = make_float3(0.0f, 0.0f, 0.0f);
    float2 sm = make_float2(0.0f, 0.0f);
    for (int yy = -3; yy <= 3; yy++)
        for (int xx = -3; xx <= 3; xx++) {
            const int qx = x + xx, qy = y + yy;
            if (!inImage(qx, qy, F.w, F.h)) continue;
            const size_t q = size_t(qy) * F.w + qx;
            const float4 oc = tempIn[q];
            const float4 om = accumMoments[q];
            const float wt = svgfWeight(cd, F.depth[q].x, cf, cn, F.normal[q], cc.x, oc.x, 10.0f);
            sw += wt;
            sc = make_float3(sc.x + wt * oc.x, sc.y + wt * oc.y, sc.z + wt * oc.z);
            sm = make_float2(sm.x + wt * om.x, sm.y + wt * om.y);
        }
    sw = fmaxf(sw, 1e-6f);
    sc = make_float3(sc.x / sw, sc.y / sw, sc.z / sw);
    sm = make_float2(sm.x / sw, sm.y / sw);
    const float var = (sm.y - sm.x * sm.x) * (4.0f / hl);
    tempOut[i] = make_float4(sc.x, sc.y, sc.z, var);
}

// Compute-ATrous iteration `it` (step 2^it): 5x5 kernel (1, 2/3, 1/6), colour scale from the 3x3-blurred variance,
// the variance channel weighted by squared weights.
__global__ void k_svgfAtrous(SvgfFrame F, const float4* __restrict__ color, float4* __restrict__ out,
                             float4* __restrict__ history, int it) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= F.w || y >= F.h) return;
    const size_t i = size_t(y) * F.w + x;
    const int step = 1 << it;
    const float kv[3] = {1.0f, 2.0f / 3.0f, 1.0f / 6.0f};
    const float kg[2][2] = {{1.0f / 4.0f, 1.0f / 8.0f}, {1.0f / 8.0f, 1.0f / 16.0f}};
    float fv = 0.0f;
    for (int yy = -1; yy <= 1; yy++)
        for (int xx = -1; xx <= 1; xx++) fv += fetchW(color, x + xx, y + yy, F.w, F.h) * kg[abs(xx)][abs(yy)];
    const float phiC = sqrtf(fmaxf(0.0f, 1e-10f + fv));
    const float4 c0 = color[i];
    const float4 cn = F.normal[i];
    const float cz = F.depth[i].x, cf = F.depth[i].w;
    float acc = kv[0] * kv[0];
    float4 sum = make_float4(c0.x * acc, c0.y * acc, c0.z * acc, c0.w * acc);
    for (int yy = -2; yy <= 2; yy++)
        for (int xx = -2; xx <= 2; xx++) {
            const int qx = x + xx * step, qy = y + yy * step;
            if (!inImage(qx, qy, F.w, F.h) || (xx == 0 && yy == 0)) continue;
            const size_t q = size_t(qy) * F.w + qx;
            const float k = kv[abs(xx)] * kv[abs(yy)];
            const float4 oc = color[q];
            const float phiD = fabsf(cf * sqrtf(float(xx * xx + yy * yy)) * float(step)) + 0.0001f;
            const float wt = svgfWeight(cz, F.depth[q].x, phiD, cn, F.normal[q], c0.x, oc.x, phiC) * k;
            sum = make_float4(sum.x + wt * oc.x, sum.y + wt * oc.y, sum.z + wt * oc.z, sum.w + wt * wt * oc.w);
            acc += wt;
        }
    const float4 r = make_float4(sum.x / acc, sum.y / acc, sum.z / acc, sum.w / (acc * acc));
    if (it == 0) history[i] = r;
    out[i] = r;
}

__global__ void k_svgfKeepHistory(SvgfFrame F, const float4* __restrict__ accumMoments, float4* __restrict__ prevNormal,
                                  float4* __restrict__ prevDepth, float4* __restrict__ momentsHistory) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= F.w || y >= F.h) return;
    const size_t i = size_t(y) * F.w + x;
    prevNormal[i] = F.normal[i];
    prevDepth[i] = F.depth[i];
    momentsHistory[i] = accumMoments[i];
}

__global__ void k_copy4(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

hipError_t launchSvgf(const SvgfInputs& in, const SvgfState& st, int iterations, float4** result,
                      hipStream_t stream) {
    const dim3 block(16, 16);
    const dim3 grid((in.width + 15) / 16, (in.height + 15) / 16);
    SvgfFrame F;
    F.noisy = in.noisy;
    F.flow = in.flow;
    F.normal = in.normal;
    F.depth = in.depth;
    F.prevNormal = st.prevNormal;
    F.prevDepth = st.prevDepth;
    F.colorHistory = st.colorHistory;
    F.momentsHistory = st.momentsHistory;
    F.w = in.width;
    F.h = in.height;
    F.allowedZ = 0.002f;   // SVGF.hpp:110-114
    F.allowedN = 0.02f;
    hipLaunchKernelGGL(k_svgfReproject, grid, block, 0, stream, F, st.accumMoments, st.tempA);
    hipLaunchKernelGGL(k_svgfFilterMoments, grid, block, 0, stream, F, st.accumMoments, st.tempA, st.tempB);
    const size_t n = size_t(in.width) * in.height;
    float4* cur = st.tempB;
    if (iterations < 1) {
        hipLaunchKernelGGL(k_copy4, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, cur, st.colorHistory, n);
    } else {
        float4* bufs[2] = {st.tempA, st.pong};
        for (int it = 0; it < iterations; it++) {
            float4* dst = bufs[it & 1];
            hipLaunchKernelGGL(k_svgfAtrous, grid, block, 0, stream, F, cur, dst, st.colorHistory, it);
            cur = dst;
        }
    }
    hipLaunchKernelGGL(k_svgfKeepHistory, grid, block, 0, stream, F, st.accumMoments, st.prevNormal, st.prevDepth,
                       st.momentsHistory);
    *result = cur;
    return hipGetLastError();
}

}  // namespace crvpt
