// pw_probe.hip -- vector-issue ceilings of the device, measured.
//
// The pair kernels are bound by vector-instruction issue, and what a CU can
// issue depends on the form of the instructions: on gfx950 an FP32 multiply
// or FMA whose operands are all vector registers issues at a higher rate
// than one that takes a scalar-register operand (the scalar-fed kernels'
// source values), and FP64 FMAs at a lower one.  These kernels run long
// chains of one instruction form on every CU, each lane with 8 independent
// accumulators (no dependence stalls at any occupancy used), so that the
// roofline of a pair kernel can be stated against the issue rate the device
// reaches for the instructions that kernel is made of, next to the published
// peak.  No counterpart in the reference: measurement only.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "pawan_gpu.h"
#include "pw_ctx.h"

namespace pw {
namespace {

constexpr int kProbeTB = 256;   // 4 waves per block
constexpr int kChains = 8;      // independent accumulators per lane
constexpr int kUnroll = 16;     // steps per loop trip: 128 instructions

// form 0: FP32 x = fma(x, a, b), a and b per lane (vector registers)
// form 1: FP32 x = fma(x, s, b), s a kernel argument (scalar register)
// form 2: FP64 x = fma(x, a, b), a and b per lane
// form 3: FP32 x = x * a, a per lane
template <int Form>
__global__ __launch_bounds__(kProbeTB) void k_issue(int iters, float s, float* __restrict__ out) {
  using T = typename std::conditional<Form == 2, double, float>::type;
  const T a = static_cast<T>(0.9999f + 1e-9f * threadIdx.x);
  const T b = static_cast<T>(1e-4f + 1e-9f * threadIdx.x);
  T x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = static_cast<T>(1 + c + threadIdx.x);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
#pragma unroll
      for (int c = 0; c < kChains; ++c) {
        if (Form == 1) x[c] = fmaf(x[c], s, b);
        else if (Form == 3) x[c] = x[c] * a;
        else x[c] = fma(x[c], a, b);
      }
  }
  T t = 0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) t += x[c];
  out[blockIdx.x * kProbeTB + threadIdx.x] = static_cast<float>(t);  // keeps the chains live
}

}  // namespace
}  // namespace pw

extern "C" int pw_valu_ceiling(int form, int waves_per_cu, double* rate, int* cus_out) {
  using namespace pw;
  if (rate == nullptr) return pwi::fail(PW_ERR_ARG, "pw_valu_ceiling: null rate");
  if (form < 0 || form > 3) return pwi::fail(PW_ERR_ARG, "pw_valu_ceiling: form must be 0..3");
  if (waves_per_cu < 4 || waves_per_cu > 32 || waves_per_cu % 4)
    return pwi::fail(PW_ERR_ARG, "pw_valu_ceiling: waves_per_cu must be 4..32, a multiple of 4");
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                              hipSuccess || cus <= 0)
    return pwi::fail(PW_ERR_HIP, "pw_valu_ceiling: no device");
  const int blocks = cus * (waves_per_cu / 4);
  constexpr int kIters = 4000;
  float* out = nullptr;
  if (hipMalloc(&out, static_cast<size_t>(blocks) * kProbeTB * sizeof(float)) != hipSuccess)
    return pwi::fail(PW_ERR_HIP, "pw_valu_ceiling: hipMalloc failed");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = PW_OK;
  float ms = 0.0f;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = PW_ERR_HIP;
  for (int rep = 0; rep < 2 && rc == PW_OK; ++rep) {  // the first launch warms up
    if (hipEventRecord(e0, nullptr) != hipSuccess) { rc = PW_ERR_HIP; break; }
    switch (form) {
      case 0: hipLaunchKernelGGL(k_issue<0>, dim3(blocks), dim3(kProbeTB), 0, nullptr, kIters, 0.9999f, out); break;
      case 1: hipLaunchKernelGGL(k_issue<1>, dim3(blocks), dim3(kProbeTB), 0, nullptr, kIters, 0.9999f, out); break;
      case 2: hipLaunchKernelGGL(k_issue<2>, dim3(blocks), dim3(kProbeTB), 0, nullptr, kIters, 0.9999f, out); break;
      default: hipLaunchKernelGGL(k_issue<3>, dim3(blocks), dim3(kProbeTB), 0, nullptr, kIters, 0.9999f, out);
    }
    if (hipGetLastError() != hipSuccess || hipEventRecord(e1, nullptr) != hipSuccess ||
        hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = PW_ERR_HIP;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(out);
  if (rc != PW_OK || !(ms > 0.0f)) return pwi::fail(PW_ERR_HIP, "pw_valu_ceiling: launch or timing failed");
  // wave instructions per CU per nanosecond
  const double winst = static_cast<double>(blocks) * (kProbeTB / 64) * kIters * kUnroll * kChains;
  *rate = winst / cus / (static_cast<double>(ms) * 1e6);
  if (cus_out) *cus_out = cus;
  return PW_OK;
}
