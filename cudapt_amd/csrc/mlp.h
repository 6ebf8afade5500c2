// Per-pixel error MLP 9 -> 64 -> 64 -> 1: parameter layout and launch interface
// of mlp.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cpt {

// Flat FP32 parameter vector (row-major, input index first):
//   W1[9][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3
constexpr int kMlpIn = 9;
constexpr int kMlpHidden = 64;
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpIn * kMlpHidden;      // 576
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHidden;                // 640
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpHidden * kMlpHidden;   // 4736
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHidden;                // 4800
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpHidden;                // 4864
constexpr int kMlpParams = kMlpOffB3 + 1;                        // 4865
static_assert(kMlpParams == 4865, "parameter count");

// Operands in the fragment order the MFMA instructions read, BF16 matrices
// (one 8- or 16-byte load per lane and fragment), FP32 vectors.
constexpr int kMlpW1Frag = 4 * 64 * 4;       // W1^T: 4 hidden blocks x 64 lanes x 4 (K = 16, features 9..15 zero)
constexpr int kMlpW2Frag = 4 * 2 * 64 * 8;   // W2 or W2^T: 4 row blocks x 2 k-blocks x 64 lanes x 8
struct MlpDevWeights {
  uint16_t* w1f;  // layer 1 A operands
  uint16_t* w2tf; // layer 2 A operands (W2^T, k-permuted)
  uint16_t* w2f;  // backward A operands (W2, k-permuted)
  float* b1;      // [64]
  float* b2;      // [64]
  float* w3;      // [64]
  float* b3;      // [1]
};

struct MlpState {
  float* w;       // FP32 master parameters
  float* m;       // Adam first moment
  float* v;       // Adam second moment
  float* grad;    // last summed gradient
  float* loss;    // last batch loss (sum of squared residuals)
  MlpDevWeights dev;
};

// Forward + backward of n samples (x: n x 9 row-major, t: n targets) on
// mlp_train_groups(n) workgroups; writes one gradient vector per workgroup
// into `part` (groups x 4865) and one loss partial (sum of squared residuals)
// per workgroup.  The loss is mean squared error.
constexpr uint32_t kMlpMaxTrainGroups = 256;
uint32_t mlp_train_groups(uint32_t n);
hipError_t launch_mlp_train(const MlpState& st, const float* x, const float* t, uint32_t n, float* part,
                            float* loss_part, hipStream_t s);
// Sums the nwg partials into st.grad / st.loss in a fixed order, and when
// `apply` is set performs Adam step `step` (1-based); refreshes the operands.
hipError_t launch_mlp_adam(const MlpState& st, const float* part, uint32_t nwg, const float* loss_part, float lr,
                           int step, int apply, hipStream_t s);
hipError_t launch_mlp_infer(const MlpState& st, const float* x, uint32_t n, float* y, hipStream_t s);
// Rebuilds the operand pack from st.w.
hipError_t launch_mlp_pack(const MlpState& st, hipStream_t s);
size_t mlp_lds_bytes();

// Renderer-side kernels (images W*H x 3 FP32, n = W*H pixels):
// the 9 features of each pixel, the log per-sample-variance training target
// against a converged image, and the allocation signal from model outputs y.
struct MlpFeatureArgs {
  const float* result;      // running mean
  const float* filtered;    // cross-bilateral output
  const float* err;         // filter residual
  const float* sq;          // running mean of squared samples
  const float* sq_f;        // the same, cross-bilateral filtered
  const float* gpos;
  const float* gnrm;
  const float* gdif;
  const uint32_t* sample_n;
  float cam[3];
};
hipError_t launch_mlp_features(const MlpFeatureArgs& a, uint32_t n, float* x, hipStream_t s);
hipError_t launch_mlp_target(const float* result, const float* ref, const uint32_t* sample_n, uint32_t n, float clamp,
                             float* t, hipStream_t s);
// signal 0: nd / (nd + N) * exp(y) / N (the filter-residual estimate's shape);
// signal 1: exp(y) / N (the pixel's expected squared error); signal 2:
// exp(y) / N^2; signal 3: sqrt(exp(y) / N * plain), the geometric mean of the
// learned estimate and the plain error map (filter residual or empirical).
hipError_t launch_mlp_err(const float* y, const uint32_t* sample_n, uint32_t nd, uint32_t signal, uint32_t n,
                          const float* plain, float* e,
                          hipStream_t s);

}  // namespace cpt
