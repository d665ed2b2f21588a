// Batched render queue: many requests of different volumes, resolutions and sample counts traced by one persistent
// launch (no reference counterpart; the reference renders one frame per call, Module.cpp:286-317).
//
// Work decomposition: request r owns the batch units [unitBegin_r, unitBegin_r + groups_r * W_r * H_r); unit
// unitBegin_r + g * (W_r*H_r) + p traces samples [g*G, min(spp_r, (g+1)*G)) of pixel p, one after another on the
// same lane, and leaves one record (the sum of its samples' radiance). The units of one group are neighbouring
// pixels, so a wave starts on coherent camera rays as in the single-frame pool. Lanes claim units exactly like the
// single-frame pool (ballot, one atomic per wave, lane rank by popcount); a lane whose unit has samples left starts
// the next one in place, without claiming. The request of a unit is found by binary search over the requests' first
// units; a lane reads its request's scene constants and volume through the device request table.
//
// Every path draws the same counter-based random numbers as in a single-frame render of the same request (pixel,
// frame number, seed offset) and runs the same device code, so a batch with G = 1 reproduces the single-frame images
// up to the last bits (the compiler forms fused multiply-adds differently in the two kernels); larger G also
// regroups the floating-point sum over samples.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "path_tracer.hpp"

namespace crvpt {

// A request's volume, read through the device request table: dense or NanoVDB by P.sparse.
struct BatchGrid {
    const BatchRequest* R;
    using Cache = LeafCache;
    CRVPT_DEV float lookupWorldJ(const SceneParams& P, V3 pos, float jx, float jy, float jz, LeafCache& c) const {
        if (R->P.sparse) return NanoAccess{R->nano}.lookupWorldJ(P, pos, jx, jy, jz, c);
        return DenseAccess{R->dense}.lookupWorldJ(P, pos, jx, jy, jz);
    }
};
// Batches write no feature maps.
CRVPT_DEV V3 cloudFiniteDifference(const BatchGrid&, const SceneParams&, const PathState&, V3) { return v3(0.0f); }

// Last request r with begins[r] <= q (begins ascending, begins[0] = 0).
CRVPT_DEV int findRequest(const uint64_t* __restrict__ begins, int n, uint64_t q) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begins[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

#ifndef CRVPT_BATCH_CLAIM_PARTS
#define CRVPT_BATCH_CLAIM_PARTS 1
#endif
// Unit partitions: the batch's units can be split into this many contiguous ranges (consecutive requests), each with
// its own claim counter; a workgroup first drains range blockIdx.x % parts, then the following ones. One range (a
// single counter over all units) measured fastest on the 64-request workload (8 ranges: 4 % slower), so it is the
// default.
constexpr uint32_t kBatchParts = CRVPT_BATCH_CLAIM_PARTS;

CRVPT_DEV uint32_t unitPartBegin(uint32_t totalUnits, uint32_t part) {
    return uint32_t(uint64_t(totalUnits) * part / kBatchParts);
}

template <int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(batchPoolWaves<MODE>()))) void k_batchPool(
        const BatchRequest* __restrict__ reqs, const uint64_t* __restrict__ unitBegins, int nreq, uint32_t totalUnits,
        uint32_t samplesPerUnit, const uint32_t* __restrict__ frameOffset, float4* __restrict__ records,
        uint32_t* __restrict__ counter, int stepsPerRefill) {
    const int lane = __lane_id();
    const uint32_t frameOff = *frameOffset;
    PathState st;
    st.flags = 0;
    st.unit = 0;
    st.pix = 0;
    st.taps = 0;
    int req = 0;
    uint32_t left = 0, sample = 0;
    V3 sum = v3(0.0f);
    bool exhausted = false;
    Walker W;
    W.svReads = 0;
    LeafCache gc;
    // A finished sample adds to its unit's sum; the unit's last sample stores the record.
    auto finishSample = [&]() {
        sum += st.color;
        st.flags = 0;
        if (left == 0u) records[st.unit] = make_float4(sum.x, sum.y, sum.z, 0.0f);
    };
    uint32_t part = kBatchParts == 1u ? 0u : blockIdx.x % kBatchParts;  // wave-uniform
    uint32_t partsLeft = kBatchParts;
    uint32_t begin = unitPartBegin(totalUnits, part), count = unitPartBegin(totalUnits, part + 1u) - begin;
    while (true) {
        while (true) {
            const bool claim = !(st.flags & PF_ALIVE) && left == 0u && !exhausted;
            const uint64_t claimMask = __ballot(claim);
            if (claimMask != 0ull) {
                const int leader = __ffsll((unsigned long long)claimMask) - 1;
                const uint32_t want = uint32_t(__popcll(claimMask));
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(&counter[part], want);
                base = __shfl(base, leader);
                const bool drained = base + want >= count;
                const uint32_t partStart = begin, partCount = count;
                if (drained) {
                    // this range is drained: the next one, until every range is
                    if (--partsLeft == 0u) {
                        exhausted = true;
                    } else {
                        part = part + 1u == kBatchParts ? 0u : part + 1u;
                        begin = unitPartBegin(totalUnits, part);
                        count = unitPartBegin(totalUnits, part + 1u) - begin;
                    }
                }
                if (claim) {
                    const uint32_t c = base + uint32_t(__popcll(claimMask & ((1ull << lane) - 1ull)));
                    const uint32_t u = partStart + c;
                    if (c < partCount) {
                        const int r = findRequest(unitBegins, nreq, u);
                        if (r != req) {
                            gc = LeafCache();  // the memo holds byte offsets into the previous request's grid
                            req = r;
                        }
                        const BatchRequest& R = reqs[req];
                        const uint32_t npix = uint32_t(R.P.width) * uint32_t(R.P.height);
                        const uint32_t local = u - uint32_t(unitBegins[req]);
                        const uint32_t g = local / npix;
                        st.unit = u;
                        st.pix = local - g * npix;
                        sample = g * samplesPerUnit;
                        left = min(samplesPerUnit, R.spp - sample);
                        sum = v3(0.0f);
                    }
                }
            }
            if (!(st.flags & PF_ALIVE) && left > 0u) {
                const BatchRequest& R = reqs[req];
                st.frame = R.P.frameBase + frameOff + sample;
                sample++;
                left--;
                if (!startPathAt<MODE>(st, W, R.P, R.S, nullptr, R.P.invViewProj)) finishSample();
            }
            if (__ballot(!(st.flags & PF_ALIVE) && (left > 0u || !exhausted)) == 0ull) break;
        }
        if (__ballot((st.flags & PF_ALIVE) != 0) == 0ull) break;  // every lane is exhausted
        for (int k = 0; k < stepsPerRefill; k++) {
            if ((st.flags & (PF_ALIVE | PF_PENDING)) == PF_ALIVE) {
                const BatchRequest& R = reqs[req];
                if (!stepPath<MODE, BatchGrid>(st, W, R.P, R.S, BatchGrid{&R}, gc, nullptr)) finishSample();
            }
            const uint64_t pend = __ballot((st.flags & PF_PENDING) != 0);
            if (pend != 0ull && (__popcll(pend) >= kResolveLanes || k == stepsPerRefill - 1)) {
                if (st.flags & PF_PENDING) {
                    const BatchRequest& R = reqs[req];
                    if (!resolvePending<MODE, BatchGrid>(st, W, R.P, R.S, BatchGrid{&R}, nullptr)) finishSample();
                }
            }
        }
    }
}

// Per pixel of every request: the sum of its units' records in group order, divided by spp (the single-frame
// accumulation's order and rounding when every unit holds one sample), written as [C, H, W] with alpha 1.
__global__ void k_batchResolve(const BatchRequest* __restrict__ reqs, const uint64_t* __restrict__ pixBegins,
                               int nreq, uint64_t totalPix, const float4* __restrict__ records) {
    const uint64_t q = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= totalPix) return;
    const int r = findRequest(pixBegins, nreq, q);
    const BatchRequest& R = reqs[r];
    const uint32_t npix = uint32_t(R.P.width) * uint32_t(R.P.height);
    const uint32_t p = uint32_t(q - pixBegins[r]);
    // FP64 sums over the units (see k_accumulate: float sums of many equal samples drift)
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (uint32_t g = 0; g < R.groups; g++) {
        const float4 v = records[R.unitBegin + uint64_t(g) * npix + p];
        sx += double(v.x);
        sy += double(v.y);
        sz += double(v.z);
    }
    const double inv = 1.0 / double(R.spp);
    float* o = R.out;
    o[p] = float(sx * inv);
    o[npix + p] = float(sy * inv);
    o[2 * size_t(npix) + p] = float(sz * inv);
    if (R.channels == 4) o[3 * size_t(npix) + p] = 1.0f;
}

__global__ void k_setU32(uint32_t* p, uint32_t v) {
    if (threadIdx.x == 0) *p = v;
}

template <int MODE>
static const void* batchKernel() {
    return (const void*)k_batchPool<MODE>;
}

static const void* batchPoolKernel(int mode) {
    switch (mode) {
        case VPT_RATIO_TRACKING: return batchKernel<VPT_RATIO_TRACKING>();
        case VPT_NEXT_EVENT_TRACKING: return batchKernel<VPT_NEXT_EVENT_TRACKING>();
        case VPT_SPECTRAL_DELTA_TRACKING: return batchKernel<VPT_SPECTRAL_DELTA_TRACKING>();
        case VPT_NEXT_EVENT_TRACKING_SPECTRAL: return batchKernel<VPT_NEXT_EVENT_TRACKING_SPECTRAL>();
        case VPT_DECOMPOSITION_TRACKING: return batchKernel<VPT_DECOMPOSITION_TRACKING>();
        case VPT_RESIDUAL_RATIO_TRACKING: return batchKernel<VPT_RESIDUAL_RATIO_TRACKING>();
        default: return batchKernel<VPT_DELTA_TRACKING>();
    }
}

int batchOccupancyBlocksPerCU(int mode, int blockSize) {
    int blocks = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, batchPoolKernel(mode), blockSize, 0);
    if (e != hipSuccess || blocks <= 0) return 1;
    return blocks;
}

hipError_t launchSetU32(uint32_t* p, uint32_t v, hipStream_t stream) {
    hipLaunchKernelGGL(k_setU32, dim3(1), dim3(64), 0, stream, p, v);
    return hipGetLastError();
}

hipError_t launchBatch(int mode, const BatchRequest* reqs, const uint64_t* unitBegins, const uint64_t* pixBegins,
                       int nreq, uint32_t totalUnits, uint64_t totalPix, const BatchLaunch& cfg,
                       const uint32_t* frameOffset, float4* records, uint32_t* counter, hipStream_t stream) {
    if (cfg.blockSize <= 0 || cfg.blockSize > 256 || cfg.blockSize % 64 != 0 || cfg.numBlocks == 0 ||
        cfg.samplesPerUnit == 0 || cfg.stepsPerRefill <= 0)
        return hipErrorInvalidValue;
    for (uint32_t p = 0; p < kBatchParts; p++) hipLaunchKernelGGL(k_setU32, dim3(1), dim3(64), 0, stream, counter + p, 0u);
    void* args[] = {(void*)&reqs,    (void*)&unitBegins, (void*)&nreq,    (void*)&totalUnits,
                    (void*)&cfg.samplesPerUnit, (void*)&frameOffset, (void*)&records, (void*)&counter,
                    (void*)&cfg.stepsPerRefill};
    hipError_t e = hipLaunchKernel(batchPoolKernel(mode), dim3(cfg.numBlocks), dim3(cfg.blockSize), args, 0, stream);
    if (e != hipSuccess) return e;
    const uint64_t blocks = (totalPix + 255) / 256;
    hipLaunchKernelGGL(k_batchResolve, dim3(unsigned(blocks)), dim3(256), 0, stream, reqs, pixBegins, nreq, totalPix,
                       (const float4*)records);
    return hipGetLastError();
}

}  // namespace crvpt
