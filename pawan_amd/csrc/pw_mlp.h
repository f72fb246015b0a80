// pw_mlp.h -- the far-field network: parameter layout, packed inference
// weights, training kernels and sample generation.  Internal to
// libpawan_gpu.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "pw_internal.h"

namespace pw {

// FP32 master parameters, in this order (W[k][j] = input k -> unit j):
//   W1[16][128] b1[128] W2[128][128] b2[128] W3[128][128] b3[128]
//   W4[128][12] b4[12]
constexpr int kMlpIn = 16, kMlpHid = 128, kMlpOut = 12;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpIn * kMlpHid;
constexpr int kOffW2 = kOffB1 + kMlpHid;
constexpr int kOffB2 = kOffW2 + kMlpHid * kMlpHid;
constexpr int kOffW3 = kOffB2 + kMlpHid;
constexpr int kOffB3 = kOffW3 + kMlpHid * kMlpHid;
constexpr int kOffW4 = kOffB3 + kMlpHid;
constexpr int kOffB4 = kOffW4 + kMlpHid * kMlpOut;
constexpr int kMlpParams = kOffB4 + kMlpOut;  // 36,748

// Inference weights: BF16 operand fragments in the lane order of the
// matrix-core instructions (layer 1: 8 tiles x 64 lanes x 4; layers 2, 3:
// 8 tiles x 4 K blocks x 64 lanes x 8; layer 4: 4 K blocks x 64 lanes x 8),
// then the biases b1 | b2 | b3 | b4 (padded to 16) in FP32.
constexpr int kMlpFragElems = 8 * 64 * 4 + 2 * (8 * 4 * 64 * 8) + 4 * 64 * 8;
constexpr int kMlpFragBytes = 2 * kMlpFragElems;  // 73,728
constexpr int kMlpBiasLen = 3 * kMlpHid + 16;

struct MlpWeights {
  const uint16_t* frag;  // kMlpFragElems BF16
  const float* bias;     // kMlpBiasLen
};

// FP32 -> BF16, round to nearest even
uint16_t bf16_bits(float x);
// master parameters -> fragments + biases
void pack_mlp(const float* params, std::vector<uint16_t>& frag, std::vector<float>& bias);
// weights and biases uniform in +-1/sqrt(fan_in), from a seed
void init_mlp(uint64_t seed, float* params);

// Per-sample side data written with the training samples (doubles).
constexpr int kSampleAux = 12;

// Inference over given scaled inputs (n x 16) -> n x 12.
hipError_t launch_mlp_forward(const float* X, int n, float* Y, const MlpWeights& W, hipStream_t st);
// Training samples from the far lists (spec: n x {leaf, slot, target}).
hipError_t launch_ff_samples(const Rec64* sorted, const NodeRec* nodes, const NodeQuad* quad, const int* far_list,
                             int cap, const int* spec, int nsamples, float* X, float* Y, double* aux,
                             hipStream_t st);

// Training (FP32).  One call = loss and gradient of
//   L = (1/n) sum_s sum_o (f(x_s)_o - y_so)^2
// over the samples idx[0..n) of (X, Y): per-block partial gradients into
// gpart (blocks x kMlpParams), per-block loss sums into lpart, then the
// ordered reduction into grad (and *loss_dev = L).
constexpr int kTrainRows = 32;  // samples per block
int train_blocks(int n);
hipError_t launch_mlp_grad(const float* params, const float* X, const float* Y, const int* idx, int n,
                           float* gpart, float* lpart, float* grad, float* loss_dev, hipStream_t st);
// Adam step on params with the gradient in grad.
hipError_t launch_adam(float* params, const float* grad, float* m, float* v, float lr, float b1,
                       float b2, float eps, int step, hipStream_t st);

}  // namespace pw
