// pw_stage.hip -- stage arithmetic of the time integrators (FP64).
//
// The reference advances the concatenated state vector with GSL BLAS-1
// calls (src/integration/rk4.cpp:12-77, rk3.cpp:12-65,
// integration.cpp:126-136); these kernels do the same element-wise
// operations in the same order, with multiply and add kept separate
// (contraction off), on the device-resident particle records.
#include <hip/hip_runtime.h>

#include "pw_internal.h"

namespace pw {
namespace {

constexpr int kEW = 256;

__global__ __launch_bounds__(kEW) void k_stage(const Rec64* __restrict__ y,
                                               const Rate64* __restrict__ k,
                                               double c, Rec64* __restrict__ out,
                                               int n) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kEW + threadIdx.x;
  if (i >= n) return;
  const Rec64 a = y[i];
  const Rate64 r = k[i];
  Rec64 o;
  o.x = r.ux * c + a.x;
  o.y = r.uy * c + a.y;
  o.z = r.uz * c + a.z;
  o.s2 = a.s2;
  o.ax = r.wx * c + a.ax;
  o.ay = r.wy * c + a.ay;
  o.az = r.wz * c + a.az;
  o.vol = a.vol;
  out[i] = o;
}

__device__ __forceinline__ double rk4c(double y, double a1, double a4,
                                       double a2, double a3, double c6,
                                       double c3) {
#pragma clang fp contract(off)
  return y + ((a1 + a4) * c6 + (a2 + a3) * c3);
}

__global__ __launch_bounds__(kEW) void k_rk4_final(
    Rec64* __restrict__ y, const Rate64* __restrict__ k1,
    const Rate64* __restrict__ k2, const Rate64* __restrict__ k3,
    const Rate64* __restrict__ k4, double c6, double c3, int n) {
  const int i = blockIdx.x * kEW + threadIdx.x;
  if (i >= n) return;
  Rec64 a = y[i];
  const Rate64 r1 = k1[i], r2 = k2[i], r3 = k3[i], r4 = k4[i];
  a.x = rk4c(a.x, r1.ux, r4.ux, r2.ux, r3.ux, c6, c3);
  a.y = rk4c(a.y, r1.uy, r4.uy, r2.uy, r3.uy, c6, c3);
  a.z = rk4c(a.z, r1.uz, r4.uz, r2.uz, r3.uz, c6, c3);
  a.ax = rk4c(a.ax, r1.wx, r4.wx, r2.wx, r3.wx, c6, c3);
  a.ay = rk4c(a.ay, r1.wy, r4.wy, r2.wy, r3.wy, c6, c3);
  a.az = rk4c(a.az, r1.wz, r4.wz, r2.wz, r3.wz, c6, c3);
  y[i] = a;
}

__global__ __launch_bounds__(kEW) void k_rk3_mid(const Rec64* __restrict__ y,
                                                 Rate64* __restrict__ k1,
                                                 const Rate64* __restrict__ k2,
                                                 double dt, double two_dt,
                                                 Rec64* __restrict__ out, int n) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kEW + threadIdx.x;
  if (i >= n) return;
  const Rec64 a = y[i];
  Rate64 r1 = k1[i];
  const Rate64 r2 = k2[i];
  r1.ux = r1.ux * dt; r1.uy = r1.uy * dt; r1.uz = r1.uz * dt;
  r1.wx = r1.wx * dt; r1.wy = r1.wy * dt; r1.wz = r1.wz * dt;
  Rec64 o;
  o.x = (r2.ux * two_dt + a.x) - r1.ux;
  o.y = (r2.uy * two_dt + a.y) - r1.uy;
  o.z = (r2.uz * two_dt + a.z) - r1.uz;
  o.s2 = a.s2;
  o.ax = (r2.wx * two_dt + a.ax) - r1.wx;
  o.ay = (r2.wy * two_dt + a.ay) - r1.wy;
  o.az = (r2.wz * two_dt + a.az) - r1.wz;
  o.vol = a.vol;
  k1[i] = r1;
  out[i] = o;
}

__device__ __forceinline__ double rk3c(double y, double b1, double a2,
                                       double a3, double c6, double c23) {
#pragma clang fp contract(off)
  return y + ((b1 * (1.0 / 6.0) + a3 * c6) + a2 * c23);
}

__global__ __launch_bounds__(kEW) void k_rk3_final(
    Rec64* __restrict__ y, const Rate64* __restrict__ k1,
    const Rate64* __restrict__ k2, const Rate64* __restrict__ k3, double c6,
    double c23, int n) {
  const int i = blockIdx.x * kEW + threadIdx.x;
  if (i >= n) return;
  Rec64 a = y[i];
  const Rate64 r1 = k1[i], r2 = k2[i], r3 = k3[i];
  a.x = rk3c(a.x, r1.ux, r2.ux, r3.ux, c6, c23);
  a.y = rk3c(a.y, r1.uy, r2.uy, r3.uy, c6, c23);
  a.z = rk3c(a.z, r1.uz, r2.uz, r3.uz, c6, c23);
  a.ax = rk3c(a.ax, r1.wx, r2.wx, r3.wx, c6, c23);
  a.ay = rk3c(a.ay, r1.wy, r2.wy, r3.wy, c6, c23);
  a.az = rk3c(a.az, r1.wz, r2.wz, r3.wz, c6, c23);
  y[i] = a;
}

__global__ __launch_bounds__(kEW) void k_euler(Rec64* __restrict__ y,
                                               const Rate64* __restrict__ k,
                                               double dt, int n) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kEW + threadIdx.x;
  if (i >= n) return;
  Rec64 a = y[i];
  const Rate64 r = k[i];
  a.x = a.x + r.ux * dt;
  a.y = a.y + r.uy * dt;
  a.z = a.z + r.uz * dt;
  a.ax = a.ax + r.wx * dt;
  a.ay = a.ay + r.wy * dt;
  a.az = a.az + r.wz * dt;
  y[i] = a;
}

__global__ __launch_bounds__(kEW) void k_check_finite(
    const Rate64* __restrict__ k, int n, int* __restrict__ flag) {
  const int i = blockIdx.x * kEW + threadIdx.x;
  if (i >= n) return;
  const Rate64 r = k[i];
  const double s = r.ux + r.uy + r.uz + r.wx + r.wy + r.wz;
  if (!isfinite(s) || !isfinite(r.ux) || !isfinite(r.wx)) atomicOr(flag, 1);
}

__global__ void k_fill_int(int* __restrict__ p, int n, int value) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

__global__ void k_zero_doubles(double* __restrict__ p, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ __launch_bounds__(kEW) void k_add_rates(Rate64* __restrict__ acc, const Rate64* __restrict__ add,
                                                   int first, int last) {
  const int i = first + blockIdx.x * kEW + static_cast<int>(threadIdx.x);
  if (i >= last) return;
  Rate64 a = acc[i];
  const Rate64 b = add[i];
  a.ux += b.ux; a.uy += b.uy; a.uz += b.uz;
  a.wx += b.wx; a.wy += b.wy; a.wz += b.wz;
  acc[i] = a;
}

inline int ew_blocks(int n) { return (n + kEW - 1) / kEW; }

}  // namespace

hipError_t launch_zero_doubles(double* p, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_doubles, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

hipError_t launch_add_rates(Rate64* acc, const Rate64* add, int first, int last, hipStream_t st) {
  if (last <= first) return hipSuccess;
  hipLaunchKernelGGL(k_add_rates, dim3(ew_blocks(last - first)), dim3(kEW), 0, st, acc, add, first, last);
  return hipGetLastError();
}

hipError_t launch_stage(const Rec64* y, const Rate64* k, double c, Rec64* out,
                        int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_stage, dim3(ew_blocks(n)), dim3(kEW), 0, st, y, k, c, out, n);
  return hipGetLastError();
}

hipError_t launch_rk4_final(Rec64* y, const Rate64* k1, const Rate64* k2,
                            const Rate64* k3, const Rate64* k4, double c6,
                            double c3, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rk4_final, dim3(ew_blocks(n)), dim3(kEW), 0, st, y, k1, k2, k3, k4,
                     c6, c3, n);
  return hipGetLastError();
}

hipError_t launch_rk3_mid(const Rec64* y, Rate64* k1, const Rate64* k2,
                          double dt, double two_dt, Rec64* out, int n,
                          hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rk3_mid, dim3(ew_blocks(n)), dim3(kEW), 0, st, y, k1, k2, dt, two_dt,
                     out, n);
  return hipGetLastError();
}

hipError_t launch_rk3_final(Rec64* y, const Rate64* k1, const Rate64* k2,
                            const Rate64* k3, double c6, double c23, int n,
                            hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rk3_final, dim3(ew_blocks(n)), dim3(kEW), 0, st, y, k1, k2, k3, c6,
                     c23, n);
  return hipGetLastError();
}

hipError_t launch_euler(Rec64* y, const Rate64* k, double dt, int n,
                        hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_euler, dim3(ew_blocks(n)), dim3(kEW), 0, st, y, k, dt, n);
  return hipGetLastError();
}

hipError_t launch_fill_int(int* p, int n, int value, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_int, dim3((n + 63) / 64), dim3(64), 0, st, p, n, value);
  return hipGetLastError();
}

hipError_t launch_check_finite(const Rate64* k, int n, int* flag,
                               hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_check_finite, dim3(ew_blocks(n)), dim3(kEW), 0, st, k, n, flag);
  return hipGetLastError();
}

}  // namespace pw
