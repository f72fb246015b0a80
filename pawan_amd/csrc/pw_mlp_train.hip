// pw_mlp_train.hip -- training of the far-field network (FP32 master
// weights, FP32 arithmetic) and the packing of its BF16 inference
// fragments.
//
// One block trains on 32 samples: forward through the four layers with the
// pre-activations kept in LDS, then back-propagation of the squared error,
// writing the block's gradient to its own row of a partial buffer; a second
// kernel sums the rows in block order (deterministic) and a third applies
// Adam.  GELU is the tanh form x sigmoid(2 z), z = sqrt(2/pi)(x + 0.044715
// x^3), the same function as the inference kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "pw_mlp.h"

namespace pw {
namespace {

constexpr int R = kTrainRows;
constexpr int H = kMlpHid;

__device__ __forceinline__ float gelu_f(float x) {
  const float z2 = -2.0f * 0.7978845608f * fmaf(0.044715f * x * x, x, x);
  return x / (1.0f + __expf(z2));
}
__device__ __forceinline__ float gelu_d(float x) {
  const float z2 = -2.0f * 0.7978845608f * fmaf(0.044715f * x * x, x, x);
  const float sg = 1.0f / (1.0f + __expf(z2));
  const float dz = 0.7978845608f * fmaf(3.0f * 0.044715f, x * x, 1.0f);
  return sg + x * sg * (1.0f - sg) * 2.0f * dz;
}

struct TrainLds {
  float x[R][kMlpIn];
  float a1[R][H], a2[R][H], a3[R][H];
  float h[R][H];
  float o[R][16];
  float dA[R][H], dB[R][H];
  float loss[4];
};

// out[s][j] = b[j] + sum_k in[s][k] W[k][j], N = 128 (thread: one j, 16 rows)
template <int K>
__device__ __forceinline__ void dense128(const float (*in)[K], const float* __restrict__ W,
                                         const float* __restrict__ b, float (*out)[H], int rows) {
  const int j = threadIdx.x & 127, s0 = threadIdx.x >> 7;
  float acc[R / 2];
#pragma unroll
  for (int i = 0; i < R / 2; ++i) acc[i] = b[j];
  for (int k = 0; k < K; ++k) {
    const float w = W[k * H + j];
#pragma unroll
    for (int i = 0; i < R / 2; ++i) acc[i] = fmaf(in[s0 + 2 * i][k], w, acc[i]);
  }
#pragma unroll
  for (int i = 0; i < R / 2; ++i)
    if (s0 + 2 * i < rows) out[s0 + 2 * i][j] = acc[i];
}

// d_prev[s][k] = gelu'(a[s][k]) sum_j d[s][j] W[k][j]  (thread: one k, 16 rows)
template <int N, int LD>
__device__ __forceinline__ void back128(const float (*d)[LD], const float* __restrict__ W,
                                        const float (*a)[H], float (*dprev)[H]) {
  const int k = threadIdx.x & 127, s0 = threadIdx.x >> 7;
  float acc[R / 2];
#pragma unroll
  for (int i = 0; i < R / 2; ++i) acc[i] = 0.0f;
  for (int j = 0; j < N; ++j) {
    const float w = W[k * N + j];
#pragma unroll
    for (int i = 0; i < R / 2; ++i) acc[i] = fmaf(d[s0 + 2 * i][j], w, acc[i]);
  }
#pragma unroll
  for (int i = 0; i < R / 2; ++i) dprev[s0 + 2 * i][k] = acc[i] * gelu_d(a[s0 + 2 * i][k]);
}

// g[k * N + j] = sum_s in[s][k] d[s][j];  g_b[j] = sum_s d[s][j]
template <int K, int N, int LDI, int LDD>
__device__ __forceinline__ void wgrad(const float (*in)[LDI], const float (*d)[LDD], float* __restrict__ gW,
                                      float* __restrict__ gb) {
  for (int e = threadIdx.x; e < K * N; e += blockDim.x) {
    const int k = e / N, j = e % N;
    float acc = 0.0f;
#pragma unroll 8
    for (int s = 0; s < R; ++s) acc = fmaf(in[s][k], d[s][j], acc);
    gW[e] = acc;
  }
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    float acc = 0.0f;
    for (int s = 0; s < R; ++s) acc += d[s][j];
    gb[j] = acc;
  }
}

__device__ __forceinline__ void gelu_rows(const float (*a)[H], float (*h)[H]) {
  for (int e = threadIdx.x; e < R * H; e += blockDim.x) h[e / H][e % H] = gelu_f(a[e / H][e % H]);
}

__global__ __launch_bounds__(256) void k_mlp_grad(const float* __restrict__ P, const float* __restrict__ X,
                                                  const float* __restrict__ Y, const int* __restrict__ idx,
                                                  int n, float* __restrict__ gpart, float* __restrict__ lpart) {
  extern __shared__ __align__(16) unsigned char smem[];
  TrainLds& s = *reinterpret_cast<TrainLds*>(smem);
  const int row0 = blockIdx.x * R;
  const int rows = min(R, n - row0);
  // inputs (rows beyond n are zero and carry zero error)
  for (int e = threadIdx.x; e < R * kMlpIn; e += blockDim.x) {
    const int r = e / kMlpIn, c = e % kMlpIn;
    s.x[r][c] = r < rows ? X[static_cast<size_t>(idx[row0 + r]) * kMlpIn + c] : 0.0f;
  }
  __syncthreads();
  dense128<kMlpIn>(s.x, P + kOffW1, P + kOffB1, s.a1, R);
  __syncthreads();
  gelu_rows(s.a1, s.h);
  __syncthreads();
  dense128<H>(s.h, P + kOffW2, P + kOffB2, s.a2, R);
  __syncthreads();
  gelu_rows(s.a2, s.h);
  __syncthreads();
  dense128<H>(s.h, P + kOffW3, P + kOffB3, s.a3, R);
  __syncthreads();
  gelu_rows(s.a3, s.h);  // h = h3
  __syncthreads();
  // output layer and its error: o[s][j] = 2 (f - y), loss rows
  float lsum = 0.0f;
  for (int e = threadIdx.x; e < R * kMlpOut; e += blockDim.x) {
    const int r = e / kMlpOut, j = e % kMlpOut;
    float acc = P[kOffB4 + j];
    for (int k = 0; k < H; ++k) acc = fmaf(s.h[r][k], P[kOffW4 + k * kMlpOut + j], acc);
    float err = 0.0f;
    if (r < rows) err = acc - Y[static_cast<size_t>(idx[row0 + r]) * kMlpOut + j];
    s.o[r][j] = 2.0f * err;
    lsum = fmaf(err, err, lsum);
  }
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
  if ((threadIdx.x & 63) == 0) s.loss[threadIdx.x >> 6] = lsum;
  __syncthreads();
  float* g = gpart + static_cast<size_t>(blockIdx.x) * kMlpParams;
  if (threadIdx.x == 0) lpart[blockIdx.x] = (s.loss[0] + s.loss[1]) + (s.loss[2] + s.loss[3]);
  wgrad<H, kMlpOut, H, 16>(s.h, s.o, g + kOffW4, g + kOffB4);
  back128<kMlpOut, 16>(s.o, P + kOffW4, s.a3, s.dA);  // delta3
  __syncthreads();
  gelu_rows(s.a2, s.h);  // h2
  __syncthreads();
  wgrad<H, H, H, H>(s.h, s.dA, g + kOffW3, g + kOffB3);
  back128<H, H>(s.dA, P + kOffW3, s.a2, s.dB);  // delta2
  __syncthreads();
  gelu_rows(s.a1, s.h);  // h1
  __syncthreads();
  wgrad<H, H, H, H>(s.h, s.dB, g + kOffW2, g + kOffB2);
  back128<H, H>(s.dB, P + kOffW2, s.a1, s.dA);  // delta1
  __syncthreads();
  wgrad<kMlpIn, H, kMlpIn, H>(s.x, s.dA, g + kOffW1, g + kOffB1);
}

// grad[p] = sum_b gpart[b][p] / n (blocks in order); loss likewise
__global__ __launch_bounds__(256) void k_grad_reduce(const float* __restrict__ gpart, const float* __restrict__ lpart,
                                                     int blocks, int n, float* __restrict__ grad,
                                                     float* __restrict__ loss) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const float inv = 1.0f / static_cast<float>(n);
  if (p < kMlpParams) {
    float acc = 0.0f;
    for (int b = 0; b < blocks; ++b) acc += gpart[static_cast<size_t>(b) * kMlpParams + p];
    grad[p] = acc * inv;
  }
  if (p == 0) {
    float acc = 0.0f;
    for (int b = 0; b < blocks; ++b) acc += lpart[b];
    *loss = acc * inv;
  }
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ P, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= kMlpParams) return;
  const float gi = g[p];
  const float mi = b1 * m[p] + (1.0f - b1) * gi;
  const float vi = b2 * v[p] + (1.0f - b2) * gi * gi;
  m[p] = mi;
  v[p] = vi;
  P[p] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

int train_blocks(int n) { return (n + R - 1) / R; }

hipError_t launch_mlp_grad(const float* params, const float* X, const float* Y, const int* idx, int n,
                           float* gpart, float* lpart, float* grad, float* loss_dev, hipStream_t st) {
  if (n <= 0) return hipErrorInvalidValue;
  const int blocks = train_blocks(n);
  static std::atomic<unsigned long long> lds_raised{0};
  hipError_t e = ensure_lds_limit(reinterpret_cast<const void*>(k_mlp_grad), sizeof(TrainLds), lds_raised);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(256), sizeof(TrainLds), st, params, X, Y, idx, n, gpart,
                     lpart);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_grad_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, gpart, lpart, blocks,
                     n, grad, loss_dev);
  return hipGetLastError();
}

hipError_t launch_adam(float* params, const float* grad, float* m, float* v, float lr, float b1, float b2,
                       float eps, int step, hipStream_t st) {
  const float c1 = 1.0f / (1.0f - std::pow(b1, static_cast<float>(step)));
  const float c2 = 1.0f / (1.0f - std::pow(b2, static_cast<float>(step)));
  hipLaunchKernelGGL(k_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, params, grad, m, v, lr, b1, b2,
                     eps, c1, c2);
  return hipGetLastError();
}

uint16_t bf16_bits(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return 0x7fc0;  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

void pack_mlp(const float* P, std::vector<uint16_t>& frag, std::vector<float>& bias) {
  frag.assign(kMlpFragElems, 0);
  bias.assign(kMlpBiasLen, 0.0f);
  size_t o = 0;
  // layer 1: tile m, lane (c, g), slot j -> W1[4g + j][16m + c]
  for (int m = 0; m < 8; ++m)
    for (int lane = 0; lane < 64; ++lane)
      for (int j = 0; j < 4; ++j) {
        const int c = lane & 15, g = lane >> 4;
        frag[o++] = bf16_bits(P[kOffW1 + (4 * g + j) * H + 16 * m + c]);
      }
  // layers 2, 3: tile m, K block b, lane, slot j -> W[in][16m + c],
  // in = 32b + 4g + j (j < 4), 32b + 16 + 4g + j - 4 (j >= 4)
  for (int layer = 0; layer < 2; ++layer) {
    const float* W = P + (layer == 0 ? kOffW2 : kOffW3);
    for (int m = 0; m < 8; ++m)
      for (int b = 0; b < 4; ++b)
        for (int lane = 0; lane < 64; ++lane)
          for (int j = 0; j < 8; ++j) {
            const int c = lane & 15, g = lane >> 4;
            const int in = 32 * b + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4);
            frag[o++] = bf16_bits(W[in * H + 16 * m + c]);
          }
  }
  // layer 4: K block b, lane, slot j -> W4[in][c] (c < 12)
  for (int b = 0; b < 4; ++b)
    for (int lane = 0; lane < 64; ++lane)
      for (int j = 0; j < 8; ++j) {
        const int c = lane & 15, g = lane >> 4;
        const int in = 32 * b + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4);
        frag[o++] = c < kMlpOut ? bf16_bits(P[kOffW4 + in * kMlpOut + c]) : 0;
      }
  for (int j = 0; j < H; ++j) {
    bias[j] = P[kOffB1 + j];
    bias[H + j] = P[kOffB2 + j];
    bias[2 * H + j] = P[kOffB3 + j];
  }
  for (int j = 0; j < kMlpOut; ++j) bias[3 * H + j] = P[kOffB4 + j];
}

void init_mlp(uint64_t seed, float* P) {
  // weights and biases of a layer uniform in +-1/sqrt(fan_in); trains
  // markedly faster on the far-field samples than Glorot with zero biases
  uint64_t s = seed;
  auto fill = [&](int off, int count, int fan_in) {
    const double lim = 1.0 / std::sqrt(static_cast<double>(fan_in));
    for (int i = 0; i < count; ++i) {
      const double u = (splitmix(s) >> 11) * 0x1.0p-53;
      P[off + i] = static_cast<float>((2.0 * u - 1.0) * lim);
    }
  };
  fill(kOffW1, kMlpIn * H, kMlpIn);
  fill(kOffB1, H, kMlpIn);
  fill(kOffW2, H * H, H);
  fill(kOffB2, H, H);
  fill(kOffW3, H * H, H);
  fill(kOffB3, H, H);
  fill(kOffW4, H * kMlpOut, H);
  fill(kOffB4, kMlpOut, H);
}

}  // namespace pw
