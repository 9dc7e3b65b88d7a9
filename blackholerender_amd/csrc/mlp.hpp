// mlp.hpp — launchers of the surrogate's data-generation, training and inference kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"
#include "mlp_device.hpp"

namespace bhs {
namespace mlp {

// Error map of a trained model: a (log r_obs, beta) cell grid over the feature square
// [-1, 1]^2 of (x1, x2); a set bit marks a cell whose validation rays all met the escape-direction
// tolerance. Step table: the largest law-S step count of the training rays per (log r_obs, x2)
// bin, kStepBins x kStepBinsB over [log 8, log 1000] x [-1, 1] (near b_crit the step count rises
// steeply with 1 / |beta - 1|).
constexpr int kMapR = 32, kMapB = 64;
constexpr int kMapWords = kMapR * kMapB / 32;
constexpr int kStepBins = 16, kStepBinsB = 16;

// Log r_obs bin of the radius margins and the step table.
__device__ __forceinline__ int r_bin(float r) {
    const int bin = (int)((log_r(r) - kLogRLo) * ((float)kStepBins / (kLogRHi - kLogRLo)));
    return bin < 0 ? 0 : (bin >= kStepBins ? kStepBins - 1 : bin);
}

// Step-table bin of a ray in r_obs bin rb with feature x2.
__device__ __forceinline__ int step_bin_rb(int rb, float x2) {
    int b = (int)((x2 + 1.0f) * (0.5f * (float)kStepBinsB));
    b = b < 0 ? 0 : (b >= kStepBinsB ? kStepBinsB - 1 : b);
    return rb * kStepBinsB + b;
}

// Step-table bin of a ray at r_obs with feature x2.
__device__ __forceinline__ int step_bin(float r, float x2) { return step_bin_rb(r_bin(r), x2); }

// Per-model constants of the surrogate render kernel.
struct SurConst {
    uint32_t serve[kMapWords];        // error map bits, cell = ir * kMapB + ib
    uint32_t steps_max[kStepBins * kStepBinsB];   // training step maximum per (r_obs, x2) bin
    float radius_margin[kStepBins];   // relative margin around the annulus edges per log r_obs bin
    float beta_eps;                   // fallback band |beta - 1| < beta_eps
    float logit_margin;               // |logit| below this: reach undecided -> fallback
    float steps_slack;                // served only if steps_max * slack < max_steps
};

// Training arguments of one Adam step.
struct TrainArgs {
    const float* X;          // n x 3 features
    const float* Y;          // n x 5 targets
    const float* M;          // n x 5 masks
    int64_t n;
    float* P;                // kParams parameters (updated in place)
    float* G;                // kParams gradient (written)
    float* m;                // Adam moments
    float* v;
    float* ws;               // train_workspace_floats(batch) floats
    int32_t batch;
    uint32_t key;            // minibatch hash key
    int32_t ident;           // 1: minibatch row b = dataset row b (no sampling)
    int32_t reg_count;       // loss normaliser of the regression terms (<= 0: batch)
    int32_t adam;            // 0: gradient only
    float lr, b1, b2, eps, c1, c2;   // c1 = 1/(1-b1^t), c2 = 1/(1-b2^t)
    float* loss_dev;         // device float receiving the minibatch loss (may be null)
};

hipError_t launch_sample_rays(int64_t n, uint32_t key, float* pos3, float* dir3, hipStream_t s);
// Features / targets / masks of n marched rays; atomically raises steps_max[bin] (log r bins).
hipError_t launch_targets(int64_t n, const float* pos3, const float* dir3, const SoA& out, float* X,
                          float* Y, float* M, uint32_t* steps_max, hipStream_t s);
hipError_t launch_init(uint32_t key, float* P, hipStream_t s);
size_t train_workspace_floats(int32_t batch);
hipError_t launch_train_step(const TrainArgs& a, hipStream_t s);

// Inference weights: bf16 A fragments of layers 2-4 in MFMA lane order, FP32 layer 1 and biases.
constexpr size_t kPackedBytes = 2 * 32768 + 4096 + 128 * 16 + 2 * 512 + 64;
hipError_t launch_pack(const float* P, void* packed, hipStream_t s);
// Raw network outputs of n feature rows (MFMA inference).
hipError_t launch_forward(const float* X, int64_t n, const void* packed, float* Y, hipStream_t s);
// Validation: escape-direction error (rad), relative crossing-radius error and error-map cell of
// each ray's surrogate decode against integrator outputs of the same rays.
hipError_t launch_validate(int64_t n, const float* pos3, const float* dir3, const MarchConst& c,
                           const SurConst& k, const SoA& ref, const void* packed, float* err_phi,
                           float* err_r, int32_t* cell, int num_cus, hipStream_t s);

// Surrogate render of one camera: every eligible ray is served by the network, the others are
// listed for the integrator (FP64 band list in sc.band_ids, FP32 list in fb_ids).
struct SurLists {
    uint32_t* fb_ids;         // FP32 fallback items decided by the decode (after the network)
    uint32_t* fb_count;
    uint32_t* early_ids;      // FP32 fallback items known before the network (outside the served
    uint32_t* early_count;    // domain, unserved cell, step budget)
    uint32_t* early_work;     // work counter of the early march
    uint32_t* lb_ids;         // band rays the decode sends back while the band march already runs
    uint32_t* lb_count;       // (latency mode): FP64 march after the network pass
    uint32_t* lb_work;        // work counter of that march
    uint32_t fb_cap;
    uint32_t* part;           // table pass: per-block statistics partials (4 words per block)
    uint32_t* near;           // table pass, large calls: near-critical flag per 256-item chunk
};
// Side stream of the early FP32 fallback march.
struct SurStreams {
    hipStream_t aux, aux2;
    hipEvent_t ev_fork, ev_join, ev_fork2, ev_join2;
    hipEvent_t ev_tab0, ev_tab1;   // the table build on aux2 (latency mode)
};
// Table mode: per camera, the network on an n2 x n3 grid over (x2, x3) at the camera's r_obs
// (node: ncam n2 n3 x 4 floats; phi: ncam x kPhiNodes floats; range, slot: per camera).
// Table mode, per camera: an n2 x n3 grid of nodes over (x2, x3) holding the network's escape
// residual, the first two crossings' u with their residuals applied and the third crossing's u of
// the exact orbit, and a kPhiNodes-point table of the exact swept angle at escape / capture over x2
// (both over the camera's x2 range).
constexpr int32_t kPhiNodes = 4096;
// Cameras of one pose (position, orientation, field of view and image size; the jitter may
// differ) share one table: slot[k] is the first camera of camera k's pose, and the table of
// slot s holds camera s's data.
struct SurTable {
    float* node;      // ncam x n2 x n3 x 4 floats
    float* phi;       // ncam x kPhiNodes floats
    float* range;     // ncam x 2 floats (x2 bounds)
    uint32_t* slot;   // ncam table slots
    int32_t n2, n3;
    float* cam;       // ncam x 16 words: the per-camera terms of the ray setup (SurCam)
};
// tab null: per-ray network inference. ev_begin / ev_mid / ev_end (may be null) bracket the
// network part and the FP32 fallback march.
hipError_t launch_surrogate(const MarchConst& c, const SurConst& k, const void* packed,
                            const SurTable* tab, const SoA& out, const Scratch& sc,
                            const SurLists& fl, const Tuning& t, int num_cus, hipStream_t s,
                            const SurStreams& ss, hipEvent_t ev_begin, hipEvent_t ev_mid,
                            hipEvent_t ev_end);

// Batch of cameras (device camera table and item prefix sums as bhs_render_batch; no FP64
// precision).
hipError_t launch_surrogate_batch(const MarchConst* table, const uint32_t* item_base,
                                  uint32_t ncam, uint32_t n_items, bool any_band,
                                  const SurConst& k, const void* packed, const SurTable* tab,
                                  const SoA& out, const Scratch& sc, const SurLists& fl,
                                  const Tuning& t, int num_cus, hipStream_t s,
                                  const SurStreams& ss, hipEvent_t ev_begin, hipEvent_t ev_mid,
                                  hipEvent_t ev_end);

}  // namespace mlp
}  // namespace bhs
