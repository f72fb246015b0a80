// pw_pair.hip -- tiled direct-sum pair kernels for gfx950.
//
// One work item = one block: up to TB targets (one per lane) against a
// contiguous source range.  Sources are staged through LDS in tiles of TB
// records, loaded with coalesced 16-byte global loads, and read back as
// wave-uniform (broadcast) LDS reads.  Every lane accumulates the rates of
// its own target in source order, so results are deterministic and need no
// atomics.
//
// The pair arithmetic is the Gaussian-kernel INTERACT of the reference
// (src/interaction/interaction_utils.h:456-501 with KERNEL,
// src/interaction/gaussian.h:47-60), evaluated from the target's side:
//   d   = x_t - x_s,  rho = |d|,  sigma^2 = (s_s^2 + s_t^2)/2
//   u_t += q (a_s x d)
//   w_t += q (a_t x a_s) + F [d . (a_t x a_s)] d
//          + (2 nu Z / sigma^2) (V_t a_s - V_s a_t)
// which is, term by term, the contribution INTERACT adds to whichever of
// the two particles is the target (the source-side update is the same
// expression with the roles swapped and the sign flipped).
#include <hip/hip_runtime.h>

#include "pw_internal.h"

namespace pw {
namespace {

constexpr int kTB = 256;  // targets per block = sources per LDS tile

// 1/(4 pi) as the reference forms it (0.25 * M_1_PI, gaussian.h:35).
constexpr double kQ = 0x1.45f306dc9c883p-4;
// (2 pi)^1.5 (the divisor pow(2.0*M_PI,1.5) of gaussian.h:22) and its
// reciprocal.
constexpr double kP15 = 0x1.f7fccdff344acp+3;
constexpr double kK = 0x1.0411e71d77959p-4;
constexpr double kSqrtHalf = 0.70710678118654752440;

// Beyond rho^2 > kFar64 sigma^2 the Gaussian terms of q and F
// (exp(-rho_bar^2/2) and erfc(rho_bar/sqrt 2)) are below 1e-17 of the
// algebraic term 1/(4 pi rho^3), i.e. under FP64 rounding:
// q = 1/(4 pi rho^3), F = -3 q / rho^2, PSE weight ~ 0.
constexpr double kFar64 = 90.0;
// The same bound for FP32 arithmetic (below 3e-8 relative).
constexpr float kFar32 = 44.0f;
// Below rho_bar^2 < kSeries32 the FP32 path evaluates q and F from
// cancellation-free positive series (see pair32 below).
constexpr float kSeries32 = 4.0f;

// ---------------------------------------------------------------------
// FP64: the reference's formula, operation by operation.
// ---------------------------------------------------------------------
struct Acc64 {
  double ux = 0, uy = 0, uz = 0, wx = 0, wy = 0, wz = 0;
};

__device__ __forceinline__ void pair64_exact(const Rec64& t, const Rec64& s,
                                             double nu, Acc64& a) {
#pragma clang fp contract(off)
  const double dx = t.x - s.x, dy = t.y - s.y, dz = t.z - s.z;
  const double rho = sqrt(dx * dx + dy * dy + dz * dz);
  const double sigma = sqrt(0.5 * (s.s2 + t.s2));
  // KERNEL (gaussian.h:53-59)
  const double rho_bar = rho / sigma;
  const double sig3 = sigma * sigma * sigma;
  const double Z = exp(-0.5 * rho_bar * rho_bar) / sig3 / kP15;
  const double phi = kQ * erf(kSqrtHalf * rho_bar) / sig3;
  const double q = (phi / rho_bar - Z) / (rho_bar * rho_bar);
  const double F = (Z - 3.0 * q) / (rho * rho);
  // VELOCITY(q, a_src, d) (interaction_utils.h:55-62)
  a.ux += q * (s.ay * dz - s.az * dy);
  a.uy += q * (s.az * dx - s.ax * dz);
  a.uz += q * (s.ax * dy - s.ay * dx);
  // VORSTRETCH (:73-105): a_trg x a_src
  const double cx = t.ay * s.az - t.az * s.ay;
  const double cy = t.az * s.ax - t.ax * s.az;
  const double cz = t.ax * s.ay - t.ay * s.ax;
  double roaxa = 0.0;
  roaxa += dx * cx;
  roaxa += dy * cy;
  roaxa += dz * cz;
  const double fr = F * roaxa;
  // DIFFUSION (:344-378)
  const double coef = Z * nu / (0.5 * sigma * sigma);
  a.wx += (q * cx + fr * dx) + coef * (t.vol * s.ax - s.vol * t.ax);
  a.wy += (q * cy + fr * dy) + coef * (t.vol * s.ay - s.vol * t.ay);
  a.wz += (q * cz + fr * dz) + coef * (t.vol * s.az - s.vol * t.az);
}

// Same quantities with reciprocal square roots in place of the divisions
// and the closed far-field form where it is exact to FP64 rounding.
__device__ __forceinline__ void pair64_fast(const Rec64& t, const Rec64& s,
                                            double nu, Acc64& a) {
  const double dx = t.x - s.x, dy = t.y - s.y, dz = t.z - s.z;
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double sig2 = 0.5 * (s.s2 + t.s2);
  const double ir = rsqrt(r2);
  const double ir2 = ir * ir;
  double q, F, coef;
  if (r2 > kFar64 * sig2) {
    q = kQ * (ir * ir2);
    F = -3.0 * q * ir2;
    coef = 0.0;
  } else {
    const double isg = rsqrt(sig2);
    const double isg3 = isg * isg * isg;
    const double rho = r2 * ir;
    const double rb = rho * isg;              // rho / sigma
    const double irb = (sig2 * isg) * ir;     // sigma / rho
    const double Z = exp(-0.5 * rb * rb) * isg3 * kK;
    const double phi = kQ * erf(kSqrtHalf * rb) * isg3;
    q = (phi * irb - Z) * (irb * irb);
    F = (Z - 3.0 * q) * ir2;
    coef = 2.0 * nu * Z * (isg * isg);
  }
  a.ux += q * (s.ay * dz - s.az * dy);
  a.uy += q * (s.az * dx - s.ax * dz);
  a.uz += q * (s.ax * dy - s.ay * dx);
  const double cx = t.ay * s.az - t.az * s.ay;
  const double cy = t.az * s.ax - t.ax * s.az;
  const double cz = t.ax * s.ay - t.ay * s.ax;
  const double fr = F * (dx * cx + dy * cy + dz * cz);
  a.wx += q * cx + fr * dx + coef * (t.vol * s.ax - s.vol * t.ax);
  a.wy += q * cy + fr * dy + coef * (t.vol * s.ay - s.vol * t.ay);
  a.wz += q * cz + fr * dz + coef * (t.vol * s.az - s.vol * t.az);
}

template <bool kExact>
__global__ __launch_bounds__(kTB) void k_pair_fp64(
    const Rec64* __restrict__ state, const WorkItem* __restrict__ items,
    Rate64* __restrict__ rates, double nu) {
  __shared__ Rec64 tile[kTB];
  const WorkItem it = items[blockIdx.x];
  const int t = it.t0 + static_cast<int>(threadIdx.x);
  const bool active = t < it.t1;
  const Rec64 me = state[active ? t : it.t0];
  Acc64 a;
  for (int base = it.s0; base < it.s1; base += kTB) {
    const int cnt = min(kTB, it.s1 - base);
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < cnt) tile[threadIdx.x] = state[base + threadIdx.x];
    __syncthreads();
    if (active) {
      for (int l = 0; l < cnt; ++l) {
        if (base + l == t) continue;  // self pair: never evaluated
        if (kExact)
          pair64_exact(me, tile[l], nu, a);
        else
          pair64_fast(me, tile[l], nu, a);
      }
    }
  }
  if (active) {
    Rate64 r;
    r.ux = a.ux; r.uy = a.uy; r.uz = a.uz;
    r.wx = a.wx; r.wy = a.wy; r.wz = a.wz;
    rates[t] = r;
  }
}

// ---------------------------------------------------------------------
// FP32 pair math with compensated accumulation; FP64 positions.
// ---------------------------------------------------------------------
// Source record in LDS: FP64 position (the displacement is formed in FP64
// and rounded once), FP32 everything else.  48 B, three 16-B LDS reads.
struct alignas(16) SrcF {
  double x, y, z;
  float s2, ax, ay, az, vol, pad;
};

constexpr float kQf = 0x1.45f306p-4f;   // 1/(4 pi)
constexpr float kKf = 0x1.0411e8p-4f;   // (2 pi)^-1.5
constexpr int kChunk = 32;              // pairs summed before compensation

// Kahan-Babuska two-sum: hi + lo += p, error-free in the (hi, lo) pair.
__device__ __forceinline__ void two_sum(float& hi, float& lo, float p) {
  const float s = hi + p;
  const float bb = s - hi;
  const float err = (hi - (s - bb)) + (p - bb);
  hi = s;
  lo += err;
}

struct Acc32 {
  float ux = 0, uy = 0, uz = 0, wx = 0, wy = 0, wz = 0;
};

struct TargetF {
  double x, y, z;
  float s2, ax, ay, az, vol;
};

// q and F of the Gaussian kernel in FP32 without cancellation:
//   rho_bar^2 = s,  e = exp(-s/2),  B(s) = sum_k s^k / (2k+5)!!
//   q = K e (1/3 + s B) / sigma^3,  F = -3 K e B / sigma^5
// (the incomplete-gamma series of erf, A&S 6.5.29, in which every term is
// positive); for 4 <= s < 44 the closed erf form, whose cancellation is
// mild there; beyond, the far-field form.
__device__ __forceinline__ void pair32(const TargetF& t, const SrcF& s,
                                       float nu, Acc32& a) {
  const float dx = static_cast<float>(t.x - s.x);
  const float dy = static_cast<float>(t.y - s.y);
  const float dz = static_cast<float>(t.z - s.z);
  const float r2 = dx * dx + dy * dy + dz * dz;
  const float sig2 = 0.5f * (s.s2 + t.s2);
  float q, F, coef;
  if (r2 > kFar32 * sig2) {
    const float ir = rsqrtf(r2);
    const float ir2 = ir * ir;
    q = kQf * (ir * ir2);
    F = -3.0f * q * ir2;
    coef = 0.0f;
  } else {
    const float isg = rsqrtf(sig2);
    const float isg2 = isg * isg;
    const float isg3 = isg2 * isg;
    const float sb = r2 * isg2;
    const float e = expf(-0.5f * sb);
    const float Ke3 = kKf * e * isg3;  // Z
    if (sb < kSeries32) {
      // B(s) = sum_{k=0}^{11} s^k / (2k+5)!!; the k = 12 term is below
      // 2e-8 of B at s = 4.
      float B = 0x1.519298p-48f;                  // 1/27!!
      B = fmaf(B, sb, 0x1.1cd3b0p-43f);           // 1/25!!
      B = fmaf(B, sb, 0x1.bd0ac4p-39f);           // 1/23!!
      B = fmaf(B, sb, 0x1.3fdfbcp-34f);           // 1/21!!
      B = fmaf(B, sb, 0x1.a3d5a8p-30f);           // 1/19!!
      B = fmaf(B, sb, 0x1.f28db6p-26f);           // 1/17!!
      B = fmaf(B, sb, 0x1.08db48p-21f);           // 1/15!!
      B = fmaf(B, sb, 0x1.f09b28p-18f);           // 1/13!!
      B = fmaf(B, sb, 0x1.937e12p-14f);           // 1/11!!
      B = fmaf(B, sb, 0x1.1566acp-10f);           // 1/9!!
      B = fmaf(B, sb, 0x1.381382p-7f);            // 1/7!!
      B = fmaf(B, sb, 0x1.111112p-4f);            // 1/5!!
      const float A = fmaf(sb, B, 0x1.555556p-2f);  // 1/3 + s B
      q = Ke3 * A;
      F = -3.0f * Ke3 * B * isg2;
    } else {
      const float ir = rsqrtf(r2);
      const float rb = r2 * ir * isg;             // rho / sigma
      const float irb = sig2 * isg * ir;          // sigma / rho
      const float phi = kQf * erff(0.70710678f * rb) * isg3;
      q = (phi * irb - Ke3) * (irb * irb);
      F = (Ke3 - 3.0f * q) * (ir * ir);
    }
    coef = 2.0f * nu * Ke3 * isg2;
  }
  a.ux += q * (s.ay * dz - s.az * dy);
  a.uy += q * (s.az * dx - s.ax * dz);
  a.uz += q * (s.ax * dy - s.ay * dx);
  const float cx = t.ay * s.az - t.az * s.ay;
  const float cy = t.az * s.ax - t.ax * s.az;
  const float cz = t.ax * s.ay - t.ay * s.ax;
  const float fr = F * (dx * cx + dy * cy + dz * cz);
  a.wx += q * cx + fr * dx + coef * (t.vol * s.ax - s.vol * t.ax);
  a.wy += q * cy + fr * dy + coef * (t.vol * s.ay - s.vol * t.ay);
  a.wz += q * cz + fr * dz + coef * (t.vol * s.az - s.vol * t.az);
}

__global__ __launch_bounds__(kTB) void k_pair_fp32c(
    const Rec64* __restrict__ state, const WorkItem* __restrict__ items,
    Rate64* __restrict__ rates, float nu) {
  __shared__ SrcF tile[kTB];
  const WorkItem it = items[blockIdx.x];
  const int t = it.t0 + static_cast<int>(threadIdx.x);
  const bool active = t < it.t1;
  const Rec64 m = state[active ? t : it.t0];
  TargetF me;
  me.x = m.x; me.y = m.y; me.z = m.z;
  me.s2 = static_cast<float>(m.s2);
  me.ax = static_cast<float>(m.ax);
  me.ay = static_cast<float>(m.ay);
  me.az = static_cast<float>(m.az);
  me.vol = static_cast<float>(m.vol);
  float hi[6] = {0, 0, 0, 0, 0, 0}, lo[6] = {0, 0, 0, 0, 0, 0};
  for (int base = it.s0; base < it.s1; base += kTB) {
    const int cnt = min(kTB, it.s1 - base);
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < cnt) {
      const Rec64 r = state[base + threadIdx.x];
      SrcF f;
      f.x = r.x; f.y = r.y; f.z = r.z;
      f.s2 = static_cast<float>(r.s2);
      f.ax = static_cast<float>(r.ax);
      f.ay = static_cast<float>(r.ay);
      f.az = static_cast<float>(r.az);
      f.vol = static_cast<float>(r.vol);
      f.pad = 0.0f;
      tile[threadIdx.x] = f;
    }
    __syncthreads();
    if (active) {
      for (int c0 = 0; c0 < cnt; c0 += kChunk) {
        const int c1 = min(cnt, c0 + kChunk);
        Acc32 p;
        for (int l = c0; l < c1; ++l) {
          if (base + l == t) continue;
          pair32(me, tile[l], nu, p);
        }
        two_sum(hi[0], lo[0], p.ux);
        two_sum(hi[1], lo[1], p.uy);
        two_sum(hi[2], lo[2], p.uz);
        two_sum(hi[3], lo[3], p.wx);
        two_sum(hi[4], lo[4], p.wy);
        two_sum(hi[5], lo[5], p.wz);
      }
    }
  }
  if (active) {
    Rate64 r;
    r.ux = static_cast<double>(hi[0]) + static_cast<double>(lo[0]);
    r.uy = static_cast<double>(hi[1]) + static_cast<double>(lo[1]);
    r.uz = static_cast<double>(hi[2]) + static_cast<double>(lo[2]);
    r.wx = static_cast<double>(hi[3]) + static_cast<double>(lo[3]);
    r.wy = static_cast<double>(hi[4]) + static_cast<double>(lo[4]);
    r.wz = static_cast<double>(hi[5]) + static_cast<double>(lo[5]);
    rates[t] = r;
  }
}

// ---------------------------------------------------------------------
// Stage arithmetic of the steppers: FP64, separate multiply and add in the
// order the reference's gsl_vector_scale / gsl_vector_add calls use.
// ---------------------------------------------------------------------
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

inline int ew_blocks(int n) { return (n + kEW - 1) / kEW; }

}  // namespace

hipError_t launch_pair(int kind, const Rec64* state, const WorkItem* items,
                       int nitems, Rate64* rates, double nu, hipStream_t st) {
  if (nitems <= 0) return hipSuccess;
  const dim3 grid(nitems), block(kTB);
  switch (kind) {
    case kPairFP64Exact:
      hipLaunchKernelGGL(k_pair_fp64<true>, grid, block, 0, st, state, items, rates, nu);
      break;
    case kPairFP64:
      hipLaunchKernelGGL(k_pair_fp64<false>, grid, block, 0, st, state, items, rates, nu);
      break;
    case kPairFP32C:
      hipLaunchKernelGGL(k_pair_fp32c, grid, block, 0, st, state, items, rates,
                         static_cast<float>(nu));
      break;
    default:
      return hipErrorInvalidValue;
  }
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
