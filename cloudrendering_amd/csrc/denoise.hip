// Edge-avoiding a-trous wavelet denoiser (EAWDenoise.glsl:161-314 compute path, EAWDenoiser.cpp:345-391 host
// loop): maxNumIterations 5x5 passes with step widths 1, 2, 4, ..., Gaussian kernel exp(-(x^2 + y^2) / 2) and
// edge-stopping weights from colour, position and normal. The reference compute path uses the constant colour
// weight exp(-1 / phiColor) (independent of the colour distance); useColorDistance = 1 selects the paper's
// exp(-|dc|^2 / phiColor) instead. Taps outside the image are clamped to the edge.
#include <hip/hip_runtime.h>

#include "render_api.hpp"

namespace crvpt {

__device__ __forceinline__ float4 clampedTap(const float4* img, int x, int y, int w, int h) {
    x = min(max(x, 0), w - 1);
    y = min(max(y, 0), h - 1);
    return img[size_t(y) * w + x];
}

__global__ __launch_bounds__(256) void k_eawPass(const float4* __restrict__ color, const float4* __restrict__ pos,
                                                 const float4* __restrict__ nrm, float4* __restrict__ out, int w,
                                                 int h, int stepWidth, float phiColor, float phiPosition,
                                                 float phiNormal, int useColorDistance) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    const float4 c0 = color[size_t(y) * w + x];
    const float4 p0 = pos ? pos[size_t(y) * w + x] : make_float4(0, 0, 0, 0);
    const float4 n0 = nrm ? nrm[size_t(y) * w + x] : make_float4(0, 0, 0, 0);
    const float constColorWeight = fminf(__expf(-1.0f / phiColor), 1.0f);
    float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float accW = 0.0f;
#pragma unroll
    for (int i = 0; i < 25; i++) {
        const int ox = i % 5 - 2, oy = i / 5 - 2;
        const float k = __expf(-float(ox * ox + oy * oy) / 2.0f);
        const int tx = x + ox * stepWidth, ty = y + oy * stepWidth;
        const float4 c = clampedTap(color, tx, ty, w, h);
        float wgt;
        if (useColorDistance) {
            const float dx = c0.x - c.x, dy = c0.y - c.y, dz = c0.z - c.z, dw = c0.w - c.w;
            wgt = fminf(__expf(-(dx * dx + dy * dy + dz * dz + dw * dw) / phiColor), 1.0f);
        } else {
            wgt = constColorWeight;
        }
        if (pos) {
            const float4 p = clampedTap(pos, tx, ty, w, h);
            const float dx = p0.x - p.x, dy = p0.y - p.y, dz = p0.z - p.z, dw = p0.w - p.w;
            wgt *= fminf(__expf(-(dx * dx + dy * dy + dz * dz + dw * dw) / phiPosition), 1.0f);
        }
        if (nrm) {
            const float4 n = clampedTap(nrm, tx, ty, w, h);
            const float dx = n0.x - n.x, dy = n0.y - n.y, dz = n0.z - n.z, dw = n0.w - n.w;
            wgt *= fminf(__expf(-(dx * dx + dy * dy + dz * dz + dw * dw) / phiNormal), 1.0f);
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
    const dim3 block(16, 16);
    const dim3 grid((p.width + 15) / 16, (p.height + 15) / 16);
    const float4* src = color;
    float4* bufs[2] = {ping, pong};
    int stepWidth = 1;
    for (int i = 0; i < p.iterations; i++) {
        float4* dst = bufs[i & 1];
        hipLaunchKernelGGL(k_eawPass, grid, block, 0, stream, src, position, normal, dst, p.width, p.height, stepWidth,
                           p.phiColor, p.phiPosition, p.phiNormal, p.useColorDistance);
        src = dst;
        stepWidth *= 2;
    }
    *result = const_cast<float4*>(src);
    return hipGetLastError();
}

}  // namespace crvpt
