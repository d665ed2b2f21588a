// Kernel-prediction denoiser network on gfx950: six 3x3 convolutions (15 -> 32 -> 32 -> 32 -> 32 -> 32 -> 25
// channels, ReLU between them) whose 25 output logits per pixel are turned by a softmax into a 5x5 filter applied to
// the noisy radiance (the per-pixel kernel of modules/kpn_module Kernel.cu:38-83 with the caller's softmax,
// PyTorchDenoiser.cpp:428-431). Training: relative-MSE loss against a high-sample target, hand-written backward
// passes and Adam.
//
// Layout: activations are NHWC bf16 with 32 channels per pixel (the 15 input features and the 25 logits are padded
// with zero channels), so one 3x3 tap of one pixel is exactly the K = 32 operand of v_mfma_f32_16x16x32_bf16.
// Convolutions are implicit GEMMs: a workgroup stages a 16x16 output tile plus its one-pixel halo in LDS; each of
// its four waves owns four rows of 16 pixels (four 16-pixel M tiles) and both 16-channel N tiles, and accumulates
// nine MFMAs (one per tap) per M/N tile pair in FP32. Weights are pre-swizzled into per-lane MFMA fragments (bf16)
// from FP32 master copies after every optimiser step.
//
// MFMA lane layout (16x16x32, bf16 in, f32 out): lane l supplies A[m = l % 16][k = 8 * (l / 16) + j] and
// B[k = 8 * (l / 16) + j][n = l % 16] for j < 8, and receives C[m = 4 * (l / 16) + i][n = l % 16] for i < 4.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "render_api.hpp"

namespace crvpt {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kC = kKpnChannels;  // 32 channels per activation pixel
constexpr int kTile = 16;         // output tile edge in pixels
constexpr int kHalo = kTile + 2;  // staged input tile edge (3x3 halo)
constexpr int kPixStride = 40;    // bf16 per staged pixel in LDS: 32 channels + 8 of padding (80 B, 16-B aligned)

enum : int { EPI_RELU_BF16 = 0, EPI_LINEAR_F32 = 1, EPI_MASK_BF16 = 2 };

__device__ __forceinline__ f32x4 mfma16(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Stages the (16 + 2)^2 input pixels around output tile (y0, x0) of image n into LDS (zero outside the image).
__device__ __forceinline__ void stageHalo(__bf16* tile, const __bf16* __restrict__ in, int n, int H, int W, int y0,
                                          int x0) {
    for (int c = threadIdx.x; c < kHalo * kHalo * 4; c += blockDim.x) {
        const int pix = c >> 2, ch = c & 3;
        const int hy = pix / kHalo, hx = pix - hy * kHalo;
        const int gy = y0 + hy - 1, gx = x0 + hx - 1;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (gy >= 0 && gy < H && gx >= 0 && gx < W)
            v = *reinterpret_cast<const uint4*>(in + ((size_t(n) * H + gy) * W + gx) * kC + ch * 8);
        *reinterpret_cast<uint4*>(tile + pix * kPixStride + ch * 8) = v;
    }
}

// One 3x3 convolution layer over [N, H, W, 32] bf16 activations (persistent over 16x16 tiles).
//   EPI_RELU_BF16   out = bf16(max(conv + bias, 0))                      hidden layers
//   EPI_LINEAR_F32  out = f32(conv + bias)                               the logits layer
//   EPI_MASK_BF16   out = bf16(conv * (mask > 0))                        backward data pass through the layer below's
//                                                                        ReLU; flip = 1 mirrors the taps and the
//                                                                        fragments hold the transposed weights
// wsw: 9 taps x 2 N tiles x 64 lanes of bf16x8 fragments.
template <int EPI>
__global__ __launch_bounds__(256) void k_conv3x3(const __bf16* __restrict__ in, const bf16x8* __restrict__ wsw,
                                                 const float* __restrict__ bias, const __bf16* __restrict__ mask,
                                                 void* __restrict__ out, int N, int H, int W, int flip) {
    __shared__ __attribute__((aligned(16))) __bf16 tile[kHalo * kHalo * kPixStride];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kb = lane >> 4;
    bf16x8 B[9][2];
#pragma unroll
    for (int t = 0; t < 9; t++) {
        B[t][0] = wsw[(t * 2 + 0) * 64 + lane];
        B[t][1] = wsw[(t * 2 + 1) * 64 + lane];
    }
    // The MFMA takes the weights as its A operand and the pixels as B, so the 16x16 result is channels x pixels:
    // lane (m, kb) holds output channels nt * 16 + kb * 4 + 0..3 of pixel m of the tile row, four consecutive
    // channels that leave in one 8-byte (bf16) or 16-byte (f32) store.
    float bco[2][4] = {};
    if (EPI != EPI_MASK_BF16)
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
#pragma unroll
            for (int i = 0; i < 4; i++) bco[nt][i] = bias[nt * 16 + kb * 4 + i];
    const int tilesX = (W + kTile - 1) / kTile, tilesY = (H + kTile - 1) / kTile;
    const int numTiles = N * tilesX * tilesY;
    for (int ti = blockIdx.x; ti < numTiles; ti += gridDim.x) {
        const int n = ti / (tilesX * tilesY);
        const int r = ti - n * tilesX * tilesY;
        const int y0 = (r / tilesX) * kTile, x0 = (r % tilesX) * kTile;
        __syncthreads();  // the previous tile's LDS reads are done
        stageHalo(tile, in, n, H, W, y0, x0);
        __syncthreads();
        f32x4 acc[4][2];
#pragma unroll
        for (int mt = 0; mt < 4; mt++) acc[mt][0] = acc[mt][1] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int t = 0; t < 9; t++) {
            const int ky = t / 3, kx = t % 3;
            const int hy = flip ? 2 - ky : ky, hx = flip ? 2 - kx : kx;
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                const int row = wave * 4 + mt;
                const bf16x8 A =
                    *reinterpret_cast<const bf16x8*>(tile + ((row + hy) * kHalo + (m + hx)) * kPixStride + kb * 8);
                acc[mt][0] = mfma16(B[t][0], A, acc[mt][0]);
                acc[mt][1] = mfma16(B[t][1], A, acc[mt][1]);
            }
        }
        const int x = x0 + m;
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            const int y = y0 + wave * 4 + mt;
            if (y >= H || x >= W) continue;
            const size_t p = (size_t(n) * H + y) * W + x;
#pragma unroll
            for (int nt = 0; nt < 2; nt++) {
                const size_t o = p * kC + nt * 16 + kb * 4;
                const f32x4 v = acc[mt][nt];
                if (EPI == EPI_LINEAR_F32) {
                    *reinterpret_cast<float4*>(reinterpret_cast<float*>(out) + o) =
                        make_float4(v[0] + bco[nt][0], v[1] + bco[nt][1], v[2] + bco[nt][2], v[3] + bco[nt][3]);
                } else {
                    __bf16 r[4];
                    if (EPI == EPI_RELU_BF16) {
#pragma unroll
                        for (int i = 0; i < 4; i++) r[i] = (__bf16)fmaxf(v[i] + bco[nt][i], 0.0f);
                    } else {
                        const uint2 mw = *reinterpret_cast<const uint2*>(mask + o);
                        __bf16 mv[4];
                        __builtin_memcpy(mv, &mw, sizeof(mw));
#pragma unroll
                        for (int i = 0; i < 4; i++) r[i] = (__bf16)((float)mv[i] > 0.0f ? v[i] : 0.0f);
                    }
                    uint2 w;
                    __builtin_memcpy(&w, r, sizeof(w));
                    *reinterpret_cast<uint2*>(reinterpret_cast<__bf16*>(out) + o) = w;
                }
            }
        }
    }
}

// Weight gradient of one layer: dW[co][ci][tap] = sum over pixels p of dZ[p][co] * X[p + offset(tap)][ci], a GEMM
// with a 9216-element output and the pixels as its reduction dimension, split over the workgroups (split-K).
// A workgroup stages per 16x16 tile the output gradient transposed (channel-major, 16-B rows of 8 pixels) and three
// copies of the input halo shifted by 0, 1 and 2 columns (channel-major rows of 16 pixels), so that every MFMA
// operand is one aligned 16-B LDS read. Wave w owns the (co, ci) tile pair (w / 2, w % 2) and nine accumulators
// (one per tap) over all tiles the workgroup visits, and stores them as the workgroup's row of partial sums:
// part[block][((tap * 4 + wave) * 4 + i) * 64 + lane] (coalesced); k_gradReduce adds the rows in a fixed order.
constexpr int kWgradPartial = 9 * 4 * 4 * 64;  // floats per workgroup row (= 32 x 32 x 9)
static_assert(kWgradPartial == kKpnLayerParams, "one partial per weight");

__global__ __launch_bounds__(256) void k_conv3x3WeightGrad(const __bf16* __restrict__ x, const __bf16* __restrict__ dz,
                                                           float* __restrict__ part, int N, int H, int W) {
    __shared__ __attribute__((aligned(16))) __bf16 dzT[kC * kTile * kTile];            // [co][256 pixels]
    __shared__ __attribute__((aligned(16))) __bf16 xs[3 * kC * kHalo * kTile];         // [kx][ci][18 rows][16 cols]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, kb = lane >> 4;
    const int mt = wave >> 1, nt = wave & 1;
    f32x4 acc[9];
#pragma unroll
    for (int t = 0; t < 9; t++) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int tilesX = (W + kTile - 1) / kTile, tilesY = (H + kTile - 1) / kTile;
    const int numTiles = N * tilesX * tilesY;
    for (int ti = blockIdx.x; ti < numTiles; ti += gridDim.x) {
        const int n = ti / (tilesX * tilesY);
        const int r = ti - n * tilesX * tilesY;
        const int y0 = (r / tilesX) * kTile, x0 = (r % tilesX) * kTile;
        __syncthreads();
        {
            // output gradient of pixel tid of the tile, transposed into dzT
            const int py = tid >> 4, px = tid & 15;
            const int gy = y0 + py, gx = x0 + px;
            const bool in = gy < H && gx < W;
            const __bf16* src = dz + ((size_t(n) * H + (in ? gy : 0)) * W + (in ? gx : 0)) * kC;
#pragma unroll
            for (int c8 = 0; c8 < 4; c8++) {
                bf16x8 v = in ? *reinterpret_cast<const bf16x8*>(src + c8 * 8) : bf16x8{};
#pragma unroll
                for (int j = 0; j < 8; j++) dzT[(c8 * 8 + j) * 256 + tid] = in ? v[j] : (__bf16)0.0f;
            }
        }
        for (int c = tid; c < kHalo * kHalo * 4; c += 256) {
            const int pix = c >> 2, c8 = c & 3;
            const int hy = pix / kHalo, hx = pix - hy * kHalo;
            const int gy = y0 + hy - 1, gx = x0 + hx - 1;
            bf16x8 v{};
            if (gy >= 0 && gy < H && gx >= 0 && gx < W)
                v = *reinterpret_cast<const bf16x8*>(x + ((size_t(n) * H + gy) * W + gx) * kC + c8 * 8);
#pragma unroll
            for (int kx = 0; kx < 3; kx++) {
                const int col = hx - kx;
                if (col < 0 || col >= kTile) continue;
#pragma unroll
                for (int j = 0; j < 8; j++) xs[((kx * kC + c8 * 8 + j) * kHalo + hy) * kTile + col] = v[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 8; ks++) {  // 32 pixels (two tile rows) per MFMA
            const bf16x8 A = *reinterpret_cast<const bf16x8*>(dzT + (mt * 16 + m) * 256 + ks * 32 + kb * 8);
            const int py = ks * 2 + (kb >> 1), px = (kb & 1) * 8;
#pragma unroll
            for (int t = 0; t < 9; t++) {
                const int ky = t / 3, kx = t % 3;
                const bf16x8 Bv =
                    *reinterpret_cast<const bf16x8*>(xs + ((kx * kC + nt * 16 + m) * kHalo + py + ky) * kTile + px);
                acc[t] = mfma16(A, Bv, acc[t]);
            }
        }
    }
    float* row = part + size_t(blockIdx.x) * kWgradPartial;
#pragma unroll
    for (int t = 0; t < 9; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) row[((t * 4 + wave) * 4 + i) * 64 + lane] = acc[t][i];
}

// Bias gradient partials: part[block][c] = sum over the block's pixels of dZ[p][c]. Thread (pixel slot, 8-channel
// group); block reduction in LDS.
__global__ __launch_bounds__(256) void k_biasGrad(const __bf16* __restrict__ dz, float* __restrict__ part, size_t P) {
    __shared__ float red[64][kC + 1];
    const int tid = threadIdx.x, c8 = tid & 3, slot = tid >> 2;
    float s[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (size_t p = size_t(blockIdx.x) * 64 + slot; p < P; p += size_t(gridDim.x) * 64) {
        const bf16x8 v = *reinterpret_cast<const bf16x8*>(dz + p * kC + c8 * 8);
#pragma unroll
        for (int j = 0; j < 8; j++) s[j] += (float)v[j];
    }
#pragma unroll
    for (int j = 0; j < 8; j++) red[slot][c8 * 8 + j] = s[j];
    __syncthreads();
    if (tid < kC) {
        float t = 0.0f;
        for (int k = 0; k < 64; k++) t += red[k][tid];
        part[size_t(blockIdx.x) * kC + tid] = t;
    }
}

// Sums the partial rows of k_conv3x3WeightGrad (wRows rows) and k_biasGrad (bRows rows) in a fixed order and adds
// them to dW [co][ci][tap] and db [co]. Workgroups 0..143 own 64 weight partial columns each, workgroup 144 the 32
// bias columns; each thread sums every 4th (weights) or 8th (biases) row with 8 loads in flight, then LDS combines
// the row groups.
constexpr int kGradReduceBlocks = kWgradPartial / 64 + 1;

__global__ __launch_bounds__(256) void k_gradReduce(const float* __restrict__ wPart, int wRows,
                                                    const float* __restrict__ bPart, int bRows, float* __restrict__ dW,
                                                    float* __restrict__ db) {
    __shared__ float red[256];
    const int tid = threadIdx.x;
    const bool bias = blockIdx.x == kGradReduceBlocks - 1;
    const int cols = bias ? kC : 64, groups = 256 / cols;
    const int col = tid % cols, g = tid / cols;
    const int stride = bias ? kC : kWgradPartial;
    const float* src = bias ? bPart + col : wPart + blockIdx.x * 64 + col;
    const int rows = bias ? bRows : wRows;
    float s8[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    int r = g;
    for (; r + 7 * groups < rows; r += 8 * groups) {
#pragma unroll
        for (int j = 0; j < 8; j++) s8[j] += src[size_t(r + j * groups) * stride];
    }
    for (; r < rows; r += groups) s8[0] += src[size_t(r) * stride];
    red[tid] = ((s8[0] + s8[1]) + (s8[2] + s8[3])) + ((s8[4] + s8[5]) + (s8[6] + s8[7]));
    __syncthreads();
    if (g != 0) return;
    float t = 0.0f;
    for (int k = 0; k < groups; k++) t += red[k * cols + col];
    if (bias) {
        db[col] += t;
        return;
    }
    // partial column e = ((tap * 4 + wave) * 4 + i) * 64 + lane -> (co, ci, tap) as k_conv3x3WeightGrad's lanes hold it
    const int e = blockIdx.x * 64 + col;
    const int lane = e & 63, i = (e >> 6) & 3, wave = (e >> 8) & 3, tap = e >> 10;
    const int co = (wave >> 1) * 16 + (lane >> 4) * 4 + i, ci = (wave & 1) * 16 + (lane & 15);
    dW[(co * kC + ci) * 9 + tap] += t;
}

// Input features [N, 15, H, W] f32 -> [N, H, W, 32] bf16 (channels 15..31 zero). Radiance (0-2) and radiance variance
// (11-13) enter as log(1 + max(v, 0)). The screen-space depth gradient (9-10) is A / sqrt(1 - A^2) of the averaged
// view-space normal component A (Clouds.glsl:296): +-inf where the normal lies on a view axis and NaN where rounding
// puts |A| above 1, so it is clamped to +-kGradientClamp. Any other non-finite input enters as 0: one NaN would
// otherwise spread through the 3x3 receptive fields of all six layers.
constexpr float kGradientClamp = 16.0f;

__global__ void k_packFeatures(const float* __restrict__ f, __bf16* __restrict__ out, int N, int H, int W) {
    const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t plane = size_t(H) * W;
    if (p >= size_t(N) * plane) return;
    const size_t n = p / plane, q = p - n * plane;
    bf16x8 v[4];
#pragma unroll
    for (int c = 0; c < kC; c++) {
        float x = 0.0f;
        if (c < kKpnInputs) {
            x = f[(n * kKpnInputs + c) * plane + q];
            if (c == 9 || c == 10) {
                x = x != x ? 0.0f : fminf(fmaxf(x, -kGradientClamp), kGradientClamp);
            } else {
                if (!__builtin_isfinite(x)) x = 0.0f;
                if (c <= 2 || (c >= 11 && c <= 13)) x = log1pf(fmaxf(x, 0.0f));
            }
        }
        v[c >> 3][c & 7] = (__bf16)x;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) reinterpret_cast<bf16x8*>(out + p * kC)[k] = v[k];
}

// Softmax over the first 25 logits of every pixel, 5x5 filter of the noisy radiance [N, 3, H, W] (zero outside the
// image), written to pred [N, 3, H, W]. With a target: adds sum((pred - t)^2 / (t^2 + 0.01)) / (3 N H W) to *loss and
// writes the logits' gradient dZ [N, H, W, 32] bf16 (zero in the padding channels).
__global__ __launch_bounds__(256) void k_kpnHead(const float* __restrict__ logits, const float* __restrict__ noisy,
                                                 const float* __restrict__ target, float* __restrict__ pred,
                                                 __bf16* __restrict__ dz, float* __restrict__ loss, int N, int H,
                                                 int W) {
    __shared__ float red[256];
    const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t plane = size_t(H) * W;
    float lossPart = 0.0f;
    if (p < size_t(N) * plane) {
        const size_t n = p / plane, q = p - n * plane;
        const int y = int(q / W), x = int(q - size_t(y) * W);
        float w[kKpnTaps];
        float mx = -3.0e38f;
#pragma unroll
        for (int k = 0; k < kKpnTaps; k++) {
            w[k] = logits[p * kC + k];
            mx = fmaxf(mx, w[k]);
        }
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < kKpnTaps; k++) {
            w[k] = __expf(w[k] - mx);
            s += w[k];
        }
        const float inv = 1.0f / s;
#pragma unroll
        for (int k = 0; k < kKpnTaps; k++) w[k] *= inv;
        float tap[3][kKpnTaps];
        float o[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int k = 0; k < kKpnTaps; k++) {
            const int yy = y + k / 5 - 2, xx = x + k % 5 - 2;
            const bool in = yy >= 0 && yy < H && xx >= 0 && xx < W;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                tap[c][k] = in ? noisy[(n * 3 + c) * plane + size_t(yy) * W + xx] : 0.0f;
                o[c] += w[k] * tap[c][k];
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) pred[(n * 3 + c) * plane + q] = o[c];
        if (target) {
            const float scale = 1.0f / (3.0f * float(size_t(N) * plane));
            float g[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float t = target[(n * 3 + c) * plane + q];
                const float d = o[c] - t;
                const float den = 1.0f / (t * t + 0.01f);
                lossPart += d * d * den;
                g[c] = 2.0f * d * den * scale;
            }
            lossPart *= scale;
            float gw[kKpnTaps];
            float dot = 0.0f;
#pragma unroll
            for (int k = 0; k < kKpnTaps; k++) {
                gw[k] = g[0] * tap[0][k] + g[1] * tap[1][k] + g[2] * tap[2][k];
                dot += w[k] * gw[k];
            }
            bf16x8 v[4];
#pragma unroll
            for (int k = 0; k < kC; k++) v[k >> 3][k & 7] = (__bf16)(k < kKpnTaps ? w[k] * (gw[k] - dot) : 0.0f);
#pragma unroll
            for (int k = 0; k < 4; k++) reinterpret_cast<bf16x8*>(dz + p * kC)[k] = v[k];
        }
    }
    if (loss) {
        red[threadIdx.x] = lossPart;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) atomicAdd(loss, red[0]);
    }
}

// Adam on the FP32 master parameters (bias-corrected, as torch.optim.Adam without weight decay).
__global__ void k_adam(float* __restrict__ param, const float* __restrict__ grad, float* __restrict__ m1,
                       float* __restrict__ m2, int n, float lr, float beta1, float beta2, float eps, float bc1,
                       float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = grad[i];
    const float a = beta1 * m1[i] + (1.0f - beta1) * g;
    const float b = beta2 * m2[i] + (1.0f - beta2) * g * g;
    m1[i] = a;
    m2[i] = b;
    param[i] -= lr * (a / bc1) / (sqrtf(b / bc2) + eps);
}

// Master weights [layer][co][ci][tap] f32 -> per-lane MFMA fragments: forward B[k = ci][n = co] and, for the backward
// data pass, transposed B[k = co][n = ci]; [layer][tap][N tile][lane][8] bf16 each.
__global__ void k_swizzleWeights(const float* __restrict__ w, bf16x8* __restrict__ fwd, bf16x8* __restrict__ bwd,
                                 int layers) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (layer, tap, nt, lane)
    if (i >= layers * 9 * 2 * 64) return;
    const int lane = i & 63, nt = (i >> 6) & 1, t = (i >> 7) % 9, layer = i / (9 * 128);
    const float* wl = w + size_t(layer) * kKpnLayerParams;
    bf16x8 f, b;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int kk = (lane >> 4) * 8 + j, nn = nt * 16 + (lane & 15);
        f[j] = (__bf16)wl[(nn * kC + kk) * 9 + t];  // co = nn, ci = kk
        b[j] = (__bf16)wl[(kk * kC + nn) * 9 + t];  // co = kk, ci = nn
    }
    fwd[i] = f;
    bwd[i] = b;
}

__global__ void k_zeroF32(float* p, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

static unsigned blocksFor(size_t n, int per = 256) { return unsigned((n + per - 1) / per); }

hipError_t launchKpnConv(int epilogue, const void* in, const void* wsw, const float* bias, const void* mask, void* out,
                         int N, int H, int W, int flip, int blocks, hipStream_t stream) {
    const int tiles = N * ((H + kTile - 1) / kTile) * ((W + kTile - 1) / kTile);
    const dim3 grid(unsigned(tiles < blocks ? tiles : blocks)), block(256);
    const __bf16* i = reinterpret_cast<const __bf16*>(in);
    const bf16x8* w = reinterpret_cast<const bf16x8*>(wsw);
    const __bf16* mk = reinterpret_cast<const __bf16*>(mask);
    if (epilogue == EPI_RELU_BF16)
        hipLaunchKernelGGL(k_conv3x3<EPI_RELU_BF16>, grid, block, 0, stream, i, w, bias, mk, out, N, H, W, flip);
    else if (epilogue == EPI_LINEAR_F32)
        hipLaunchKernelGGL(k_conv3x3<EPI_LINEAR_F32>, grid, block, 0, stream, i, w, bias, mk, out, N, H, W, flip);
    else
        hipLaunchKernelGGL(k_conv3x3<EPI_MASK_BF16>, grid, block, 0, stream, i, w, bias, mk, out, N, H, W, flip);
    return hipGetLastError();
}

size_t kpnGradWorkFloats(int blocks) { return size_t(blocks) * (kWgradPartial + kC); }

hipError_t launchKpnWeightGrad(const void* x, const void* dz, float* dW, float* db, int N, int H, int W, int blocks,
                               float* work, hipStream_t stream) {
    const int tiles = N * ((H + kTile - 1) / kTile) * ((W + kTile - 1) / kTile);
    const int wRows = tiles < blocks ? tiles : blocks;
    float* wPart = work;
    float* bPart = work + size_t(blocks) * kWgradPartial;
    hipLaunchKernelGGL(k_conv3x3WeightGrad, dim3(unsigned(wRows)), dim3(256), 0, stream,
                       reinterpret_cast<const __bf16*>(x), reinterpret_cast<const __bf16*>(dz), wPart, N, H, W);
    const size_t P = size_t(N) * H * W;
    const size_t bb = (P + 63) / 64;
    const int bRows = int(bb < size_t(blocks) ? bb : size_t(blocks));
    hipLaunchKernelGGL(k_biasGrad, dim3(unsigned(bRows)), dim3(256), 0, stream, reinterpret_cast<const __bf16*>(dz),
                       bPart, P);
    hipLaunchKernelGGL(k_gradReduce, dim3(kGradReduceBlocks), dim3(256), 0, stream, wPart, wRows, bPart, bRows, dW,
                       db);
    return hipGetLastError();
}

hipError_t launchKpnPack(const float* features, void* out, int N, int H, int W, hipStream_t stream) {
    const size_t P = size_t(N) * H * W;
    hipLaunchKernelGGL(k_packFeatures, dim3(blocksFor(P)), dim3(256), 0, stream, features,
                       reinterpret_cast<__bf16*>(out), N, H, W);
    return hipGetLastError();
}

hipError_t launchKpnHead(const float* logits, const float* noisy, const float* target, float* pred, void* dz,
                         float* loss, int N, int H, int W, hipStream_t stream) {
    const size_t P = size_t(N) * H * W;
    hipLaunchKernelGGL(k_kpnHead, dim3(blocksFor(P)), dim3(256), 0, stream, logits, noisy, target, pred,
                       reinterpret_cast<__bf16*>(dz), loss, N, H, W);
    return hipGetLastError();
}

hipError_t launchKpnAdam(float* param, const float* grad, float* m1, float* m2, int n, float lr, float beta1,
                         float beta2, float eps, float bc1, float bc2, hipStream_t stream) {
    hipLaunchKernelGGL(k_adam, dim3(blocksFor(size_t(n))), dim3(256), 0, stream, param, grad, m1, m2, n, lr, beta1,
                       beta2, eps, bc1, bc2);
    return hipGetLastError();
}

hipError_t launchKpnSwizzle(const float* w, void* fwd, void* bwd, int layers, hipStream_t stream) {
    hipLaunchKernelGGL(k_swizzleWeights, dim3(blocksFor(size_t(layers) * 9 * 128)), dim3(256), 0, stream, w,
                       reinterpret_cast<bf16x8*>(fwd), reinterpret_cast<bf16x8*>(bwd), layers);
    return hipGetLastError();
}

hipError_t launchKpnZero(float* p, size_t n, hipStream_t stream) {
    hipLaunchKernelGGL(k_zeroF32, dim3(blocksFor(n)), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

}  // namespace crvpt
