// Kernel-prediction denoiser application kernels (modules/kpn_module/src/Kernel.cu, caller
// src/Denoiser/PyTorchDenoiser.cpp:410-443):
//   perPixelKernel  out[n, c, y, x] = sum over the K x K window of weights[n, k, y, x] * image[n, c, y + dy, x + dx],
//                   taps outside the output extent (min of image and weight sizes) skipped (Kernel.cu:38-83);
//   invIterBase     the sigmoid-gated blend of the low-resolution prediction, the previous frame and the
//                   high-frequency detail (Kernel.cu:137-186), with the reference's weight batch stride of
//                   K*K*H*W although the fused weights hold K*K + 3 channels per batch;
//   softmax         over the first K*K channels of the fused weights, per pixel (torch::softmax(l, 1)).
#include <hip/hip_runtime.h>

#include <mutex>

#include "render_api.hpp"

namespace crvpt {

constexpr int kKpnMaxChannels = 4;

__global__ void k_perPixelKernel(const float* __restrict__ image, const float* __restrict__ weights,
                                 float* __restrict__ out, int K, int h, int w, int imgH, int imgW, int wH, int wW,
                                 int C) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int n = blockIdx.z;
    if (x >= w || y >= h) return;
    float acc[kKpnMaxChannels] = {0.0f, 0.0f, 0.0f, 0.0f};
    const size_t wPlane = size_t(wW) * wH, iPlane = size_t(imgW) * imgH;
    const float* wb = weights + size_t(n) * wPlane * K * K;
    const float* ib = image + size_t(n) * iPlane * C;
    int k = 0;
    for (int dy = -K / 2; dy <= K / 2; dy++) {
        const int ky = y + dy;
        for (int dx = -K / 2; dx <= K / 2; dx++, k++) {
            const int kx = x + dx;
            if (kx < 0 || kx >= w || ky < 0 || ky >= h) continue;
            const float wt = wb[size_t(k) * wPlane + size_t(y) * wW + x];
#pragma unroll
            for (int c = 0; c < kKpnMaxChannels; c++)
                if (c < C) acc[c] += wt * ib[size_t(c) * iPlane + size_t(ky) * imgW + kx];
        }
    }
    float* ob = out + size_t(n) * size_t(w) * h * C;
#pragma unroll
    for (int c = 0; c < kKpnMaxChannels; c++)
        if (c < C) ob[size_t(c) * size_t(w) * h + size_t(y) * w + x] = acc[c];
}

__global__ void k_invIterBase(const float* __restrict__ hiX, const float* __restrict__ loPred,
                              const float* __restrict__ loX, const float* __restrict__ prev,
                              const float* __restrict__ weights, float* __restrict__ out, int usePrevious, int K, int h,
                              int w, int imgH, int imgW, int loH, int loW, int wH, int wW, int C) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int n = blockIdx.z;
    const int lx = x / 2, ly = y / 2;
    if (x >= w || y >= h) return;
    if (lx >= loW || ly >= loH) {
        // outside the low-resolution image the reference leaves its zero-initialised result (Kernel.cu:240)
        for (int c = 0; c < C; c++)
            out[size_t(n) * size_t(w) * h * C + size_t(c) * size_t(w) * h + size_t(y) * w + x] = 0.0f;
        return;
    }
    const int kd = K * K;
    const size_t wPlane = size_t(wW) * wH;
    const float* wb = weights + size_t(n) * wPlane * kd;  // the reference's stride (K*K planes per batch)
    for (int c = 0; c < C; c++) {
        float lfW = wb[size_t(kd) * wPlane + size_t(y) * wW + x];
        float hfW = wb[size_t(kd + 1) * wPlane + size_t(y) * wW + x];
        float prW = wb[size_t(kd + 2) * wPlane + size_t(y) * wW + x];
        lfW = 1.0f / (1.0f + expf(-lfW));
        hfW = 1.0f / (1.0f + expf(-hfW));
        prW = 1.0f / (1.0f + expf(-prW));
        const size_t hi = size_t(n) * size_t(imgW) * imgH * C + size_t(c) * size_t(imgW) * imgH + size_t(y) * imgW + x;
        const size_t lo = size_t(n) * size_t(loW) * loH * C + size_t(c) * size_t(loW) * loH + size_t(ly) * loW + lx;
        const float xHi = hiX[hi], xLo = loX[lo], pLo = loPred[lo];
        float lf = lfW * pLo + (1.0f - lfW) * xLo;
        // the reference blends with a double literal: (1.0 - prev_weight) and the sum are evaluated in double
        if (usePrevious) lf = float(double(lf) * (1.0 - double(prW)) + double(prev[hi] * prW));
        out[size_t(n) * size_t(w) * h * C + size_t(c) * size_t(w) * h + size_t(y) * w + x] =
            hfW * (xHi - xLo + lf) + (1.0f - hfW) * lf;
    }
}

// Softmax over channels [0, K*K) of a [N, Ct, H, W] tensor into a [N, K*K, H, W] tensor (max-subtracted).
__global__ void k_softmaxChannels(const float* __restrict__ in, float* __restrict__ out, int Ct, int kk, int H, int W) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int n = blockIdx.z;
    if (x >= W || y >= H) return;
    const size_t plane = size_t(W) * H, px = size_t(y) * W + x;
    const float* ib = in + size_t(n) * plane * Ct;
    float* ob = out + size_t(n) * plane * kk;
    float m = -3.402823466e38f;
    for (int k = 0; k < kk; k++) m = fmaxf(m, ib[size_t(k) * plane + px]);
    float s = 0.0f;
    for (int k = 0; k < kk; k++) s += expf(ib[size_t(k) * plane + px] - m);
    const float inv = 1.0f / s;
    for (int k = 0; k < kk; k++) ob[size_t(k) * plane + px] = expf(ib[size_t(k) * plane + px] - m) * inv;
}

hipError_t launchPerPixelKernel(const float* image, const float* weights, float* out, int N, int C, int imgH,
                                int imgW, int wH, int wW, int K, hipStream_t stream) {
    const int h = imgH < wH ? imgH : wH, w = imgW < wW ? imgW : wW;
    const dim3 block(16, 16, 1), grid((w + 15) / 16, (h + 15) / 16, N);
    hipLaunchKernelGGL(k_perPixelKernel, grid, block, 0, stream, image, weights, out, K, h, w, imgH, imgW, wH, wW, C);
    return hipGetLastError();
}

hipError_t launchInvIterBase(const float* hiX, const float* loPred, const float* loX, const float* prev,
                             const float* weights, float* out, bool usePrevious, int N, int C, int imgH, int imgW,
                             int loH, int loW, int wH, int wW, int K, hipStream_t stream) {
    const int h = imgH < wH ? imgH : wH, w = imgW < wW ? imgW : wW;
    const dim3 block(16, 16, 1), grid((w + 15) / 16, (h + 15) / 16, N);
    hipLaunchKernelGGL(k_invIterBase, grid, block, 0, stream, hiX, loPred, loX, prev, weights, out, int(usePrevious),
                       K, h, w, imgH, imgW, loH, loW, wH, wW, C);
    return hipGetLastError();
}

hipError_t launchSoftmaxChannels(const float* in, float* out, int N, int Ct, int kk, int H, int W,
                                 hipStream_t stream) {
    const dim3 block(16, 16, 1), grid((W + 15) / 16, (H + 15) / 16, N);
    hipLaunchKernelGGL(k_softmaxChannels, grid, block, 0, stream, in, out, Ct, kk, H, W);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------------------
// Inverse-iteration apply without allocation: the blended base image and the normalised weights live in this file's
// static device array instead of renderer-owned buffers. Each GPU holds its own copy of the array; its address on
// a device is taken once (hipGetSymbolAddress with that device current) and kept per device. Calls on one GPU share
// the array, so calls that overlap in time on one GPU race; every call overwrites the part it uses completely, so
// nothing needs clearing between calls.
// ------------------------------------------------------------------------------------------------------------
static __device__ __align__(256) float scratch[kKpnNoAllocScratchFloats];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratchMutex;
float* g_scratchPtr[kMaxScratchDevices] = {};
}  // namespace

hipError_t kpnScratchSetup(float** ptr) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratchPtr[dev] = static_cast<float*>(p);
    }
    if (ptr) *ptr = g_scratchPtr[dev];
    return hipSuccess;
}

hipError_t launchInvIterNoAlloc(const float* hiX, const float* loPred, const float* loX, const float* prev,
                                const float* weights, float* out, bool usePrevious, int N, int C, int imgH, int imgW,
                                int loH, int loW, int Ct, int wH, int wW, int K, hipStream_t stream) {
    float* base = nullptr;
    hipError_t e = kpnScratchSetup(&base);
    if (e != hipSuccess) return e;
    const int h = imgH < wH ? imgH : wH, w = imgW < wW ? imgW : wW;
    float* soft = base + size_t(N) * C * h * w;
    e = launchInvIterBase(hiX, loPred, loX, prev, weights, base, usePrevious, N, C, imgH, imgW, loH, loW, wH, wW, K,
                          stream);
    if (e != hipSuccess) return e;
    e = launchSoftmaxChannels(weights, soft, N, Ct, K * K, wH, wW, stream);
    if (e != hipSuccess) return e;
    return launchPerPixelKernel(base, soft, out, N, C, h, w, wH, wW, K, stream);
}

}  // namespace crvpt
