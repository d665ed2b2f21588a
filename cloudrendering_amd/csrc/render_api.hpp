// Host-callable launchers of the gfx950 kernels (defined in render.hip, used by the context in api.cpp).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include "vpt_types.h"

namespace crvpt {

// Per-pixel running state of the feature maps (Clouds.glsl:146-383), float4 per pixel each.
struct FeatureAcc {
    float4* firstX;      // First X: (x, y, z, coverage)
    float4* firstW;      // First W: (w, pdf_w) of the last sample
    float4* normal;      // Normal: running mean of the finite-difference normal
    float4* cloudOnly;   // Cloud Only
    float4* background;  // Background
    float4* albedo;      // Albedo
    float4* depth;       // (mean depth, mean squared depth, Depth Far mean, Depth (fwidth))
    float4* dens;        // (mean density * 1e-3, its mean square, Depth (nabla) x, y)
    float4* flow;        // (Flow x, y of the last sample, Reprojected UV)
    float4* flowRev;     // (Flow Reverse x, y of the last sample, 0, 0)
    float4* radianceSq;  // running mean of the squared sample radiance per channel (the Radiance Variance map)
};
constexpr int kFeatureAccArrays = 11;
// Feature map index of the per-pixel sample variance of the radiance: an extension after the reference's
// FeatureMapTypeVpt values (0..20), for denoisers that weight by the noise level.
constexpr int kFeatureMapRadianceVariance = 21;

struct LaunchConfig {
    uint32_t numBlocks;      // persistent workgroups of 256 threads
    int32_t stepsPerRefill;  // walker events per lane between two refill ballots (the compaction interval)
};

// Launch shape of the batched queue (the autotuner's parameters).
struct BatchLaunch {
    uint32_t numBlocks;       // persistent workgroups
    int32_t blockSize;        // threads per workgroup: 64, 128, 192 or 256
    int32_t stepsPerRefill;   // compaction interval: walker steps per lane between two refill ballots
    uint32_t samplesPerUnit;  // paths per claimed unit (consecutive samples of one pixel on one lane)
};

// Kernel-prediction denoiser network (kpn_net.hip): six 3x3 convolutions over 32-channel NHWC bf16 activations.
constexpr int kKpnChannels = 32;                       // activation channels (inputs and logits zero-padded)
constexpr int kKpnInputs = 15;                         // input feature channels
constexpr int kKpnTaps = 25;                           // 5x5 per-pixel filter taps (output logits)
constexpr int kKpnLayers = 6;
constexpr int kKpnLayerParams = kKpnChannels * kKpnChannels * 9;  // weights per layer [co][ci][ky * 3 + kx]
constexpr int kKpnParams = kKpnLayers * kKpnLayerParams + kKpnLayers * kKpnChannels;  // weights, then biases

// Seeded fBm value-noise cloud: sum over octaves o < octaves of gain^o * noise(p * baseFrequency * lacunarity^o),
// normalised by the sum of weights, optionally shaped by an anvil height profile, then smoothstep(edge0, edge1, .).
struct FbmParams {
    uint32_t seed;
    int32_t octaves;
    float lacunarity;
    float gain;
    float baseFrequency;
    float edge0;
    float edge1;
    int32_t heightFalloff;
};

// stats (nullable): 64-bit device counters [density taps, supervoxel majorant reads, paths], accumulated.
// Stores the launch constants to device memory from a kernel argument (stream-ordered, no host staging).
hipError_t launchStoreParams(SceneParams* dst, const SceneParams& P, hipStream_t stream);
// Largest numFrames of one launchRender call at this image size (claim indices stay below 2^31).
uint32_t maxFramesPerLaunch(int width, int height);
hipError_t launchRender(const SceneParams& P, const DenseGrid* dense, const NanoGrid* nano, const SupervoxelGrid& S,
                        float4* paths, uint32_t* counter, const LaunchConfig& cfg, unsigned long long* stats,
                        float4* feat, hipStream_t stream);
// Batched queue (batch.hip): one persistent path-pool launch over every request, then one resolve launch that writes
// each request's image. unitBegins / pixBegins: nreq ascending first units / first pixels (device arrays).
hipError_t launchBatch(int mode, const BatchRequest* reqs, const uint64_t* unitBegins, const uint64_t* pixBegins,
                       int nreq, uint32_t totalUnits, uint64_t totalPix, const BatchLaunch& cfg,
                       const uint32_t* frameOffset, float4* records, uint32_t* counter, hipStream_t stream);
hipError_t launchSetU32(uint32_t* p, uint32_t v, hipStream_t stream);

// KPN network kernels (kpn_net.hip). epilogue: 0 bias + ReLU -> bf16, 1 bias -> f32, 2 x (mask > 0) -> bf16.
hipError_t launchKpnConv(int epilogue, const void* in, const void* wsw, const float* bias, const void* mask, void* out,
                         int N, int H, int W, int flip, int blocks, hipStream_t stream);
// Weight and bias gradients of one layer, added to dW / db; work: kpnGradWorkFloats(blocks) floats of partial sums.
size_t kpnGradWorkFloats(int blocks);
hipError_t launchKpnWeightGrad(const void* x, const void* dz, float* dW, float* db, int N, int H, int W, int blocks,
                               float* work, hipStream_t stream);
hipError_t launchKpnPack(const float* features, void* out, int N, int H, int W, hipStream_t stream);
hipError_t launchKpnHead(const float* logits, const float* noisy, const float* target, float* pred, void* dz,
                         float* loss, int N, int H, int W, hipStream_t stream);
hipError_t launchKpnAdam(float* param, const float* grad, float* m1, float* m2, int n, float lr, float beta1,
                         float beta2, float eps, float bc1, float bc2, hipStream_t stream);
hipError_t launchKpnSwizzle(const float* w, void* fwd, void* bwd, int layers, hipStream_t stream);
hipError_t launchKpnZero(float* p, size_t n, hipStream_t stream);
int batchOccupancyBlocksPerCU(int mode, int blockSize);
// foreground: 0 plain running mean, 1 foreground map from the records' first-event flags, 2 foreground map of a
// tracker without first events (all zero).
hipError_t launchAccumulate(const float4* paths, uint32_t npix, uint32_t frames, float4* acc, uint32_t prevFrames,
                            float* outCHW, int channels, int foreground, hipStream_t stream);
hipError_t launchWriteCHW(const float4* img, uint32_t npix, float* outCHW, int channels, hipStream_t stream);
// noAccumulation: the reference's DISABLE_ACCUMULATION maps (the depth of a ray without an event is the far plane).
hipError_t launchAccumulateFeatures(const SceneParams& P, const float4* paths, const float4* feat, uint32_t npix,
                                    uint32_t frames, uint32_t prevFrames, bool noAccumulation, const FeatureAcc& acc,
                                    hipStream_t stream);
// Writes feature map `map` (FeatureMapTypeVpt order) as C planes of H*W floats; result = the image accumulator.
hipError_t launchFeatureOut(int map, const FeatureAcc& acc, const float4* result, uint32_t npix, float* outCHW,
                            int channels, hipStream_t stream);
hipError_t launchTransmittance(const SceneParams& P, const DenseGrid* dense, const NanoGrid* nano,
                               const SupervoxelGrid& S, int estimator, const float* o, const float* d,
                               const float* tmax, int nRays, int nSamples, uint32_t seed, float* mean, float* m2,
                               hipStream_t stream);
// Flat 8^3-block index of a NanoVDB grid over blocks [b0, b0 + nb) (see NanoGrid::blockLeaf); *bad (cleared by the
// caller) becomes nonzero when a leaf offset cannot be encoded.
hipError_t launchBlockTable(const NanoGrid& nano, const int b0[3], const int nb[3], uint32_t* leaf, float* tile,
                            uint32_t* bad, hipStream_t stream);
hipError_t launchMajorant(const DenseGrid* dense, const NanoGrid* nano, const int dims[3], int svSize,
                          const int grid3[3], float* maj, hipStream_t stream);
hipError_t launchRefSupervoxelStats(const DenseGrid* dense, const NanoGrid* nano, const int dims[3], int svSize,
                                    const int grid3[3], int halo, float* mn, float* mx, float* avg,
                                    hipStream_t stream);
hipError_t launchFbm(float* out, int nx, int ny, int nz, const FbmParams& F, hipStream_t stream);
// Dense FP32 field of a NanoVDB grid over its index bounding box (dims = its size), x fastest.
hipError_t launchDecodeDense(const NanoGrid& nano, const int dims[3], float* out, hipStream_t stream);
hipError_t launchNormalize(const float* raw, float* out, size_t n, float vmin, float vmax, hipStream_t stream);
int occupancyBlocksPerCU(int mode, bool sparse);
hipError_t launchGridLookup(const DenseGrid* dense, const NanoGrid* nano, int n, const int32_t* ijk, float* out,
                            hipStream_t stream);
hipError_t launchNanoLookupWorld(const NanoGrid& nano, int n, const float* xyz, float* out, int32_t* ijk,
                                 hipStream_t stream);

// Denoisers (denoise.hip)
struct EawParams {
    int32_t width, height;
    int32_t iterations;
    float phiColor, phiPosition, phiNormal;
    int32_t useColorDistance;  // 0: the reference compute path's constant colour weight exp(-1/phiColor)
};
hipError_t launchEaw(const float4* color, const float4* position, const float4* normal, float4* ping, float4* pong,
                     const EawParams& p, float4** result, hipStream_t stream);

// SVGF: this frame's maps and the history carried between frames (all W*H float4).
struct SvgfInputs {
    const float4* noisy;   // non-accumulated colour
    const float4* flow;    // motion in .xy
    const float4* normal;  // .xyz
    const float4* depth;   // depth in .x, depth fwidth in .w
    int width, height;
};
struct SvgfState {
    float4* prevNormal;
    float4* prevDepth;
    float4* colorHistory;
    float4* momentsHistory;
    float4* accumMoments;
    float4* tempA;
    float4* tempB;
    float4* pong;
};
constexpr int kSvgfBuffers = 8;

// Kernel-prediction denoiser application (kpn.hip).
hipError_t launchPerPixelKernel(const float* image, const float* weights, float* out, int N, int C, int imgH,
                                int imgW, int wH, int wW, int K, hipStream_t stream);
hipError_t launchInvIterBase(const float* hiX, const float* loPred, const float* loX, const float* prev,
                             const float* weights, float* out, bool usePrevious, int N, int C, int imgH, int imgW,
                             int loH, int loW, int wH, int wW, int K, hipStream_t stream);
hipError_t launchSoftmaxChannels(const float* in, float* out, int N, int Ct, int kk, int H, int W, hipStream_t stream);
// Floats of the inverse-iteration apply's static scratch array (128 MiB): the base image (N*C*h*w) and the
// normalised weights (N*K*K*wH*wW) of one call.
constexpr size_t kKpnNoAllocScratchFloats = size_t(32) << 20;
// Address of the scratch array on the current device (looked up once per device).
hipError_t kpnScratchSetup(float** ptr);
// crvpt_kpn_inv_iter_kernel's kernels on the scratch array: no device allocation.
hipError_t launchInvIterNoAlloc(const float* hiX, const float* loPred, const float* loX, const float* prev,
                                const float* weights, float* out, bool usePrevious, int N, int C, int imgH, int imgW,
                                int loH, int loW, int Ct, int wH, int wW, int K, hipStream_t stream);
hipError_t launchSvgf(const SvgfInputs& in, const SvgfState& st, int iterations, float4** result, hipStream_t stream);

}  // namespace crvpt
