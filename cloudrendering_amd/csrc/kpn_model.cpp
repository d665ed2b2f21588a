// Kernel-prediction denoiser: parameters, Adam state, activations and the forward / backward / update sequence of
// the six-layer network in kpn_net.hip, behind the crvpt_kpn_* entry points. One model per process.
//
// Parameters (FP32 master copy, kKpnParams floats): weights [layer][co][ci][ky * 3 + kx] (32 x 32 x 9 per layer,
// torch's Conv2d weight order), then biases [layer][co]. Real channels: layer 0 reads 15 input features, layer 5
// writes 25 logits; the padding weights are zero and their gradients stay zero (zero inputs, zero logit gradients).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "host_common.hpp"
#include "render_api.hpp"

namespace crvpt {
namespace {

static_assert(CRVPT_KPN_PARAMS == kKpnParams, "parameter count of the C ABI");

constexpr size_t kFragBytes = size_t(kKpnLayers) * 9 * 2 * 64 * 16;  // bf16x8 fragments of all layers

struct KpnModel {
    DevBuf params, grads, m1, m2, wFwd, wBwd, loss;
    DevBuf act, logits, dzA, dzB, pred, gradWork;
    size_t capPixels = 0;
    int step = 0;
    bool ready = false;
    int cuCount = 256;
    ~KpnModel() {
        for (DevBuf* b : {&params, &grads, &m1, &m2, &wFwd, &wBwd, &loss, &act, &logits, &dzA, &dzB, &pred, &gradWork})
            b->release();
    }
};

KpnModel* g_kpn = nullptr;

int requireModel() {
    if (!g_kpn || !g_kpn->ready)
        return fail(CRVPT_ERR_NOT_INITIALIZED, "kpn: no model (call crvpt_kpn_init or crvpt_kpn_set_parameters)");
    return CRVPT_OK;
}

int createModel() {
    if (!g_kpn) {
        g_kpn = new KpnModel();
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) == hipSuccess) g_kpn->cuCount = prop.multiProcessorCount;
    }
    KpnModel& k = *g_kpn;
    const size_t pb = size_t(kKpnParams) * sizeof(float);
    HIP_TRY(k.params.ensure(pb));
    HIP_TRY(k.grads.ensure(pb));
    HIP_TRY(k.m1.ensure(pb));
    HIP_TRY(k.m2.ensure(pb));
    HIP_TRY(k.wFwd.ensure(kFragBytes));
    HIP_TRY(k.wBwd.ensure(kFragBytes));
    HIP_TRY(k.loss.ensure(64));
    return CRVPT_OK;
}

// Uploads host parameters, clears the Adam state and refreshes the bf16 fragments.
int loadParameters(const std::vector<float>& host) {
    if (int rc = createModel()) return rc;
    KpnModel& k = *g_kpn;
    HIP_TRY(hipMemcpy(k.params.ptr, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(launchKpnZero(k.m1.as<float>(), kKpnParams, nullptr));
    HIP_TRY(launchKpnZero(k.m2.as<float>(), kKpnParams, nullptr));
    HIP_TRY(launchKpnZero(k.grads.as<float>(), kKpnParams, nullptr));
    HIP_TRY(launchKpnSwizzle(k.params.as<float>(), k.wFwd.ptr, k.wBwd.ptr, kKpnLayers, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    k.step = 0;
    k.ready = true;
    return CRVPT_OK;
}

int ensureActivations(size_t P) {
    KpnModel& k = *g_kpn;
    if (P <= k.capPixels) return CRVPT_OK;
    const size_t bf = P * kKpnChannels * 2;
    HIP_TRY(k.act.ensure(bf * kKpnLayers));  // inputs of the six layers
    HIP_TRY(k.logits.ensure(P * kKpnChannels * sizeof(float)));
    HIP_TRY(k.dzA.ensure(bf));
    HIP_TRY(k.dzB.ensure(bf));
    HIP_TRY(k.pred.ensure(P * 3 * sizeof(float)));
    k.capPixels = P;
    return CRVPT_OK;
}

const float* biasOf(const KpnModel& k, int layer) {
    return k.params.as<float>() + size_t(kKpnLayers) * kKpnLayerParams + size_t(layer) * kKpnChannels;
}
const void* fragOf(const DevBuf& b, int layer) {
    return b.as<uint8_t>() + size_t(layer) * 9 * 2 * 64 * 16;
}
void* actOf(const KpnModel& k, int layer, size_t P) {
    return k.act.as<uint8_t>() + size_t(layer) * P * kKpnChannels * 2;
}

int checkShape(const void* a, const void* b, int n, int h, int w) {
    if (!a || !b || n <= 0 || h <= 0 || w <= 0)
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "kpn: null buffer or empty batch");
    return CRVPT_OK;
}

// Forward pass: features -> act[0..5] -> logits.
int forward(const float* features, int n, int h, int w, hipStream_t s) {
    KpnModel& k = *g_kpn;
    const size_t P = size_t(n) * h * w;
    if (int rc = ensureActivations(P)) return rc;
    const int blocks = k.cuCount * 8;
    HIP_TRY(launchKpnPack(features, actOf(k, 0, P), n, h, w, s));
    for (int l = 0; l < kKpnLayers - 1; l++)
        HIP_TRY(launchKpnConv(0, actOf(k, l, P), fragOf(k.wFwd, l), biasOf(k, l), nullptr, actOf(k, l + 1, P), n, h,
                              w, 0, blocks, s));
    HIP_TRY(launchKpnConv(1, actOf(k, kKpnLayers - 1, P), fragOf(k.wFwd, kKpnLayers - 1),
                          biasOf(k, kKpnLayers - 1), nullptr, k.logits.ptr, n, h, w, 0, blocks, s));
    return CRVPT_OK;
}

}  // namespace
}  // namespace crvpt

using namespace crvpt;

extern "C" {

int crvpt_kpn_init(uint32_t seed) {
    // He-uniform weights over the real fan-in (9 x 15 for the first layer, 9 x 32 after), zero biases.
    std::vector<float> p(kKpnParams, 0.0f);
    std::mt19937 rng(seed);
    for (int l = 0; l < kKpnLayers; l++) {
        const int cin = l == 0 ? kKpnInputs : kKpnChannels;
        const int cout = l == kKpnLayers - 1 ? kKpnTaps : kKpnChannels;
        const float bound = std::sqrt(6.0f / float(9 * cin));
        std::uniform_real_distribution<float> u(-bound, bound);
        for (int co = 0; co < cout; co++)
            for (int ci = 0; ci < cin; ci++)
                for (int t = 0; t < 9; t++) p[size_t(l) * kKpnLayerParams + (co * kKpnChannels + ci) * 9 + t] = u(rng);
    }
    return loadParameters(p);
}

int crvpt_kpn_set_parameters(const float* params, uint64_t count) {
    if (!params || count != uint64_t(kKpnParams))
        return fail(CRVPT_ERR_INVALID_ARGUMENT,
                    "kpn_set_parameters: expected " + std::to_string(kKpnParams) + " floats");
    std::vector<float> p(params, params + kKpnParams);
    // padding channels stay zero
    for (int l = 0; l < kKpnLayers; l++)
        for (int co = 0; co < kKpnChannels; co++)
            for (int ci = 0; ci < kKpnChannels; ci++) {
                const bool pad = (l == 0 && ci >= kKpnInputs) || (l == kKpnLayers - 1 && co >= kKpnTaps);
                if (pad)
                    for (int t = 0; t < 9; t++) p[size_t(l) * kKpnLayerParams + (co * kKpnChannels + ci) * 9 + t] = 0.0f;
            }
    for (int co = kKpnTaps; co < kKpnChannels; co++)
        p[size_t(kKpnLayers) * kKpnLayerParams + size_t(kKpnLayers - 1) * kKpnChannels + co] = 0.0f;
    return loadParameters(p);
}

int crvpt_kpn_get_parameters(float* out, uint64_t count) {
    if (int rc = requireModel()) return rc;
    if (!out || count < uint64_t(kKpnParams)) return fail(CRVPT_ERR_INVALID_ARGUMENT, "kpn_get_parameters: too small");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out, g_kpn->params.ptr, size_t(kKpnParams) * sizeof(float), hipMemcpyDeviceToHost));
    return CRVPT_OK;
}

int crvpt_kpn_get_gradients(float* out, uint64_t count) {
    if (int rc = requireModel()) return rc;
    if (!out || count < uint64_t(kKpnParams)) return fail(CRVPT_ERR_INVALID_ARGUMENT, "kpn_get_gradients: too small");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out, g_kpn->grads.ptr, size_t(kKpnParams) * sizeof(float), hipMemcpyDeviceToHost));
    return CRVPT_OK;
}

int crvpt_kpn_compute_gradients(const float* features, const float* noisy, const float* target, int n, int h, int w,
                                float* loss, void* stream) {
    if (int rc = requireModel()) return rc;
    if (int rc = checkShape(features, noisy, n, h, w)) return rc;
    if (!target) return fail(CRVPT_ERR_INVALID_ARGUMENT, "kpn_compute_gradients: null target");
    KpnModel& k = *g_kpn;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const size_t P = size_t(n) * h * w;
    if (int rc = forward(features, n, h, w, s)) return rc;
    HIP_TRY(launchKpnZero(k.grads.as<float>(), kKpnParams, s));
    HIP_TRY(launchKpnZero(k.loss.as<float>(), 1, s));
    HIP_TRY(launchKpnHead(k.logits.as<float>(), noisy, target, k.pred.as<float>(), k.dzA.ptr, k.loss.as<float>(), n,
                          h, w, s));
    const int blocks = k.cuCount * 2;  // gradient workgroups: one row of partial sums each
    HIP_TRY(k.gradWork.ensure(kpnGradWorkFloats(blocks) * sizeof(float)));
    void* dz = k.dzA.ptr;
    void* dzNext = k.dzB.ptr;
    float* gw = k.grads.as<float>();
    float* gb = gw + size_t(kKpnLayers) * kKpnLayerParams;
    for (int l = kKpnLayers - 1; l >= 0; l--) {
        HIP_TRY(launchKpnWeightGrad(actOf(k, l, P), dz, gw + size_t(l) * kKpnLayerParams, gb + size_t(l) * kKpnChannels,
                                    n, h, w, blocks, k.gradWork.as<float>(), s));
        if (l > 0) {
            // gradient of the layer's input, through the ReLU that produced it
            HIP_TRY(launchKpnConv(2, dz, fragOf(k.wBwd, l), nullptr, actOf(k, l, P), dzNext, n, h, w, 1,
                                  k.cuCount * 8, s));
            std::swap(dz, dzNext);
        }
    }
    if (loss) {
        HIP_TRY(hipMemcpyAsync(loss, k.loss.ptr, sizeof(float), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    return CRVPT_OK;
}

int crvpt_kpn_adam_step(float lr, void* stream) {
    if (int rc = requireModel()) return rc;
    if (!(lr >= 0.0f)) return fail(CRVPT_ERR_INVALID_ARGUMENT, "kpn_adam_step: negative learning rate");
    KpnModel& k = *g_kpn;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    k.step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.0f - std::pow(b1, float(k.step)), bc2 = 1.0f - std::pow(b2, float(k.step));
    HIP_TRY(launchKpnAdam(k.params.as<float>(), k.grads.as<float>(), k.m1.as<float>(), k.m2.as<float>(), kKpnParams,
                          lr, b1, b2, 1e-8f, bc1, bc2, s));
    HIP_TRY(launchKpnSwizzle(k.params.as<float>(), k.wFwd.ptr, k.wBwd.ptr, kKpnLayers, s));
    return CRVPT_OK;
}

int crvpt_kpn_train_step(const float* features, const float* noisy, const float* target, int n, int h, int w,
                         float lr, float* loss, void* stream) {
    if (int rc = crvpt_kpn_compute_gradients(features, noisy, target, n, h, w, loss, stream)) return rc;
    return crvpt_kpn_adam_step(lr, stream);
}

int crvpt_kpn_denoise(const float* features, const float* noisy, float* out, int n, int h, int w, void* stream) {
    if (int rc = requireModel()) return rc;
    if (int rc = checkShape(features, noisy, n, h, w)) return rc;
    if (!out) return fail(CRVPT_ERR_INVALID_ARGUMENT, "kpn_denoise: null output");
    KpnModel& k = *g_kpn;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (int rc = forward(features, n, h, w, s)) return rc;
    HIP_TRY(launchKpnHead(k.logits.as<float>(), noisy, nullptr, out, nullptr, nullptr, n, h, w, s));
    return CRVPT_OK;
}

int crvpt_kpn_get_logits(float* out, int n, int h, int w, void* stream) {
    if (int rc = requireModel()) return rc;
    KpnModel& k = *g_kpn;
    const size_t P = size_t(n) * h * w;
    if (!out || P == 0 || P > k.capPixels) return fail(CRVPT_ERR_INVALID_ARGUMENT, "kpn_get_logits: bad shape");
    HIP_TRY(hipMemcpyAsync(out, k.logits.ptr, P * kKpnChannels * sizeof(float), hipMemcpyDeviceToDevice,
                           reinterpret_cast<hipStream_t>(stream)));
    return CRVPT_OK;
}

int crvpt_kpn_release(void) {
    if (g_kpn) {
        (void)hipDeviceSynchronize();
        delete g_kpn;
        g_kpn = nullptr;
    }
    return CRVPT_OK;
}

}  // extern "C"
