// Host-side state of the path-guiding network, shared by the guiding kernels
// (guiding.hip) and the guided integrator (render.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "kp_guide.h"

struct kp_ctx;

// Segments of the two-stage gradient reduction.
#define KP_G_SEGS 16

struct GuideState {
    float* params = nullptr;   // FP32 masters, canonical layout
    float* m = nullptr;        // Adam moments
    float* v = nullptr;
    float* grad = nullptr;     // last reduced gradient
    kp::u32x4* wf = nullptr;   // bf16 fragments
    kp::u32x4* wb = nullptr;
    float* partial = nullptr;  // per-block gradient contributions [train_blocks][params]
    float* seg = nullptr;      // first-stage sums [KP_G_SEGS][params]
    uint32_t train_blocks = 256;
    uint32_t max_batch = 1u << 18;
    // training batch (also the in-render training records)
    float* in6 = nullptr;
    float* dirs = nullptr;
    float* weights = nullptr;
    uint32_t* count = nullptr;
    float* scale = nullptr;
    float4* rec_aux = nullptr;   // [max_batch][2]: {radiance before the vertex, done}, {throughput, made}
    uint32_t* rec_valid = nullptr;   // valid records of the last pass
    int step = 0;
    float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    float3 bmin{0, 0, 0}, binv{1, 1, 1};
    // guided rendering (kp_guiding_configure)
    float bsdf_fraction = 0.5f, min_alpha = 0.5f, lr = 1e-3f;
    float weight_clamp = 0.0f;   // cap on training weights, multiple of the batch mean (0: none)
    bool relative_weights = false;   // weights relative to the pixel estimate
    uint32_t train_steps = 1, guide_after = 1, max_records = 1u << 18, pass_paths = 1u << 18;
    uint32_t max_train_passes = 0;   // passes that train train_steps steps (0 = every pass)
    uint32_t late_train_steps = 0;   // steps per pass after them (0 = frozen)
    uint32_t max_guided_depth = 0;   // 0 = every depth
    bool auto_box = true;
    float box_min[3] = {0, 0, 0}, box_max[3] = {0, 0, 0};
    uint32_t train_passes = 0;   // passes whose records trained the network
};

namespace kp {
GuideNet guide_net(kp_ctx* c);
// One training step on the batch staged in GuideState::in6/dirs/weights with
// the sample count and loss scale read from device memory.
int guide_train_device(kp_ctx* c, float lr, bool apply);
// Sets count = n and scale = 1 / sum(weights[0..n)) (0 if the sum is 0) on
// the device, and counts the records with a positive weight into rec_valid.
int guide_normalise_records(kp_ctx* c, uint32_t n);
// Position box used by the network: [lo, hi] per axis.
void guide_set_box(kp_ctx* c, const float* lo, const float* hi);
}  // namespace kp
