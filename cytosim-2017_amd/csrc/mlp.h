// Per-fiber network of the learned initial guess: sizes, device state and the
// host functions of mlp.hip.
#pragma once

#include "common.h"

namespace cymg {

constexpr int kMlpPoints = 31;                 // fibers of at most this many points get a learned guess
constexpr int kMlpIn = 384;                    // 4 groups of 3*kMlpPoints features, padded to 16
constexpr int kMlpHidden = 256;
constexpr int kMlpOut = 96;                    // 3*kMlpPoints outputs, padded; targets keep the count in the last
constexpr int kMlpLayers = 4;
constexpr float kMlpTargetClip = 32.f;        // samples with a larger |target| are not trained on
constexpr long kMlpWeights = (long)kMlpIn * kMlpHidden + kMlpHidden + 2L * (kMlpHidden * kMlpHidden + kMlpHidden) +
                             (long)kMlpHidden * kMlpOut + kMlpOut;

struct MlpState {
    bool ready = false;                        // weights present
    bool step_ready = false;                   // features of the current step computed
    long adam_t = 0;
    DevBuf<float> w, m, v, g;                  // FP32 master weights, Adam moments, gradients
    DevBuf<uint16_t> wb;                       // BF16 copy used by the GEMMs
    int work_rows = 0;
    DevBuf<float> z[kMlpLayers];               // pre-activations (FP32)
    DevBuf<uint16_t> h[kMlpLayers];            // layer inputs (BF16)
    DevBuf<float> dz[kMlpLayers];              // gradients of the pre-activations
    DevBuf<uint16_t> dzb[kMlpLayers];
    DevBuf<float> dh, tgt, loss;
    DevBuf<uint16_t> fx;                       // features of the step
    DevBuf<float> fs;                          // per-fiber scale |rhs|_inf
    DevBuf<double> y, x0;                      // M^-1 b and the guess
    DevBuf<double> z1;                         // M^-1 (b - A M^-1 b): what the coupled neighbours add
    DevBuf<double> gpart;                      // guard: per-block max |b - A x0| and max |b|
    DevBuf<int> gflag;                         // guard: 1 when the guess was dropped
    int* h_gflag = nullptr;                    // pinned copy of gflag
    int guard = 1;                             // CYMG_OPT_GUESS_GUARD
    // recorded (features, targets) ring
    bool recording = false;
    uint32_t ds_cap = 0, ds_count = 0, ds_next = 0;
    DevBuf<uint16_t> ds_x;
    DevBuf<float> ds_t;
};

struct Context;
void mlp_init(Context& c, uint64_t seed);
void mlp_set_weights(Context& c, const float* w);
void mlp_get_weights(Context& c, float* w);
double mlp_check(Context& c, const float* X, const float* T, int B, float* grads);
double mlp_train(Context& c, int steps, int B, double lr, uint64_t seed);
// loss of the current weights on B recorded samples drawn with `seed` (no update)
double mlp_eval(Context& c, int B, uint64_t seed);
void mlp_prepare_step(Context& c, int precond);
const double* mlp_guess(Context& c);
bool mlp_guess_dropped(Context& c);
void mlp_record(Context& c);
void mlp_features(Context& c, float* X, float* scale);
void mlp_set_record(Context& c, uint32_t capacity);

}  // namespace cymg
