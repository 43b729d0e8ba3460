// Learned search radii: a 3-layer FP32 MLP (F -> 64 -> 64 -> 1, ReLU)
// trained on the device with hand-written forward / backward / Adam kernels
// to predict log core distances from KD-leaf statistics, and the
// ball-bounded kNN it feeds.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fh {

constexpr int kMlpHidden = 64;
constexpr int kMlpMaxF = 18;  // 2 + 2 d, d <= 8

// Parameter block layout (floats): W1[H][F], b1[H], W2[H][H], b2[H], W3[H], b3[1]
inline int mlp_param_count(int F) { return kMlpHidden * F + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden + kMlpHidden + 1; }

struct MlpParams {
    float *w;      // [P] parameters
    float *g;      // [P] gradient of the last batch
    float *m, *v;  // [P] Adam moments
    float *mu, *sd;  // [F] feature standardisation
    int F;
};

// Features of each tree position: leaf count, leaf spacing, leaf extents,
// position inside the leaf box.
hipError_t launch_mlp_features(int64_t n, int d, int levels, const float *xp, const int2 *range, const float *box,
                               float *feat, hipStream_t st);
// mean / std of the features of `count` rows listed in rows (or the first count)
hipError_t launch_mlp_standardise(int count, int F, const float *feat, const int *rows, float *mu, float *sd,
                                  hipStream_t st);
hipError_t launch_mlp_init(const MlpParams &p, uint32_t seed, hipStream_t st);
// One Adam step on the batch rows[off .. off+B) (wrapping) of the training
// set: features of position rows[i], target tgt[i].  scratch holds the
// activations and deltas of the batch: B * (F + 4 * H + 2) floats.
hipError_t launch_mlp_train_step(const MlpParams &p, int ntrain, int B, int off, const float *feat, const int *rows,
                                 const float *tgt, float *scratch, float lr, int step, hipStream_t st);
// forward pass only, all n positions: out[p] = prediction
hipError_t launch_mlp_predict(const MlpParams &p, int64_t n, const float *feat, const int *rows, float *out,
                              hipStream_t st);
// residual statistics on the training rows: stats[0] = mean, stats[1] = std
// of (target - prediction)
hipError_t launch_mlp_residual(int ntrain, const float *pred, const float *tgt, float *stats, hipStream_t st);
// training targets: 0.5 log(core_sq) of the listed positions
hipError_t launch_mlp_log_targets(int ntrain, const int *rows, const float *core_p, float *tgt, hipStream_t st);
// bound_sq[p] = exp(2 (pred[p] + mean + z * std)) (squared radius)
hipError_t launch_mlp_bounds(int64_t n, const float *pred, const float *stats, float z, float *bound_sq,
                             hipStream_t st);

}  // namespace fh
