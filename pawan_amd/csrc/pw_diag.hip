// pw_diag.hip -- FP64 conserved-quantity diagnostics on the device.
//
// __interaction::diagnose (src/interaction/interaction.cpp:64-194) has two
// kinds of sums:
//   * O(N), per wake: total vorticity sum(a), linear impulse sum(x x a)/2,
//     angular impulse sum(x x (x x a))/3 and the centroid sums
//     |a| z (x^2+y^2) and |a| (x^2+y^2)  (interaction.cpp:489-527, 135-157);
//   * O(N^2), over every unordered pair counted twice plus self terms:
//     helicity, kinetic energy and its div-free variant, enstrophy and its
//     div-free variant (interaction_utils.h:107-331; the enstrophies are
//     identically zero for the Gaussian kernel, gaussian.h:69-92, and
//     nonzero for the high-order algebraic one, highorder.h:77-106).
// Both are computed here as per-block partial sums in a fixed tree order;
// the host adds the partials in block order, so results are deterministic.
// Pair terms are symmetric in the two particles, so the ordered sum over
// j != i equals the reference's 2 * sum over i < j.
#include <hip/hip_runtime.h>

#include "pw_internal.h"

namespace pw {
namespace {

constexpr int kTB = 256;
constexpr double kPi = 3.14159265358979323846;
constexpr double k1Pi = 0.31830988618379067154;  // M_1_PI
constexpr double kQ = 0x1.45f306dc9c883p-4;      // 0.25 * M_1_PI
constexpr double kP15 = 0x1.f7fccdff344acp+3;    // pow(2 pi, 1.5)
constexpr double kSqrtHalf = 0.70710678118654752440;

// Sum of v over the block, exact same tree every call: butterfly within
// each 64-lane wavefront, then the four wave totals in wave order.
template <int M>
__device__ __forceinline__ void block_sum(double (&v)[M], double* red) {
  for (int off = 32; off > 0; off >>= 1)
    for (int m = 0; m < M; ++m) v[m] += __shfl_xor(v[m], off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
    for (int m = 0; m < M; ++m) red[wave * M + m] = v[m];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int m = 0; m < M; ++m) {
      double s = red[m];
      for (int w = 1; w < kTB / 64; ++w) s += red[w * M + m];
      v[m] = s;
    }
  }
}

// Linear sums over blocks of up to kTB particles of one wake.
// items[b] = {first, last, wake, 0}; out[b * 11 + m].
__global__ __launch_bounds__(kTB) void k_diag_linear(
    const Rec64* __restrict__ y, const WorkItem* __restrict__ items,
    double* __restrict__ out) {
  __shared__ double red[(kTB / 64) * 11];
  const WorkItem it = items[blockIdx.x];
  const int i = it.t0 + static_cast<int>(threadIdx.x);
  double v[11];
  for (int m = 0; m < 11; ++m) v[m] = 0.0;
  if (i < it.t1) {
    const Rec64 p = y[i];
    // x x a
    const double rx = p.y * p.az - p.z * p.ay;
    const double ry = p.z * p.ax - p.x * p.az;
    const double rz = p.x * p.ay - p.y * p.ax;
    v[0] = p.ax; v[1] = p.ay; v[2] = p.az;
    v[3] = rx; v[4] = ry; v[5] = rz;
    // x x (x x a)
    v[6] = p.y * rz - p.z * ry;
    v[7] = p.z * rx - p.x * rz;
    v[8] = p.x * ry - p.y * rx;
    const double oy = sqrt(p.ax * p.ax + p.ay * p.ay + p.az * p.az);
    const double rr = p.x * p.x + p.y * p.y;
    v[9] = oy * p.z * rr;
    v[10] = oy * rr;
  }
  block_sum<11>(v, red);
  if (threadIdx.x == 0)
    for (int m = 0; m < 11; ++m) out[blockIdx.x * 11 + m] = v[m];
}

// QSIG (gaussian.h:30-37) with the reference's operation order.
__device__ __forceinline__ double qsig(double rho, double sigma) {
#pragma clang fp contract(off)
  const double rho_bar = rho / sigma;
  const double sig3 = sigma * sigma * sigma;
  const double Z = exp(-0.5 * rho_bar * rho_bar) / sig3 / kP15;
  const double phi = kQ * erf(kSqrtHalf * rho_bar) / sig3;
  return (phi / rho_bar - Z) / (rho_bar * rho_bar);
}

// QSIG of the high-order algebraic kernel (highorder.h:31-39), same
// operation order, pow(x, 2.5) as x^2 sqrt(x).
__device__ __forceinline__ double qsig_ho(double rho, double sigma) {
#pragma clang fp contract(off)
  const double rho_bar = rho / sigma;
  const double rho_bar2 = rho_bar * rho_bar;
  const double rho3 = rho * rho * rho;
  const double b = rho_bar2 + 1.0;
  return 0.25 * k1Pi * rho_bar2 * rho_bar * (rho_bar2 + 2.5) / (b * b * sqrt(b)) / rho3;
}

// Per-target rows of helicity, KE, KEF (and, for the high-order kernel,
// enstrophy and its div-free variant) over one work item's sources; the
// block total goes to out[kDiagQuad * b].  radius is needed for the self
// terms.
template <bool HighOrder>
__global__ __launch_bounds__(kTB) void k_diag_pairs(
    const Rec64* __restrict__ y, const double* __restrict__ radius,
    const WorkItem* __restrict__ items, double* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ Rec64 tile[kTB];
  __shared__ double red[(kTB / 64) * kDiagQuad];
  const WorkItem it = items[blockIdx.x];
  const int t = it.t0 + static_cast<int>(threadIdx.x);
  const bool active = t < it.t1;
  const Rec64 me = y[active ? t : it.t0];
  double h = 0.0, ke = 0.0, kef = 0.0, en = 0.0, enf = 0.0;
  for (int base = it.s0; base < it.s1; base += kTB) {
    const int cnt = min(kTB, it.s1 - base);
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < cnt) tile[threadIdx.x] = y[base + threadIdx.x];
    __syncthreads();
    if (active) {
      for (int l = 0; l < cnt; ++l) {
        if (base + l == t) continue;
        const Rec64 s = tile[l];
        // x12 = x_J - x_I with I = target, J = source
        const double x0 = s.x - me.x, x1 = s.y - me.y, x2 = s.z - me.z;
        double rho2 = 0.0;
        rho2 += x0 * x0;
        rho2 += x1 * x1;
        rho2 += x2 * x2;
        // HELICITY (interaction_utils.h:214-242)
        const double rho = sqrt(rho2);
        const double sigma = sqrt(0.5 * (me.s2 + s.s2));
        const double q = HighOrder ? qsig_ho(rho, sigma) : qsig(rho, sigma);
        const double c0 = me.ay * s.az - me.az * s.ay;
        const double c1 = me.az * s.ax - me.ax * s.az;
        const double c2 = me.ax * s.ay - me.ay * s.ax;
        double roaxa = 0.0;
        roaxa += x0 * c0;
        roaxa += x1 * c1;
        roaxa += x2 * c2;
        h += q * roaxa;
        // KINETICENERGY / KINETICENERGYF (:278-331)
        const double sigma2 = 0.5 * (me.s2 + s.s2);
        double a1a2 = 0.0, x12a1 = 0.0, x12a2 = 0.0;
        a1a2 += me.ax * s.ax;
        a1a2 += me.ay * s.ay;
        a1a2 += me.az * s.az;
        x12a1 += x0 * me.ax;
        x12a1 += x1 * me.ay;
        x12a1 += x2 * me.az;
        x12a2 += x0 * s.ax;
        x12a2 += x1 * s.ay;
        x12a2 += x2 * s.az;
        const double b = rho2 + sigma2;
        const double p15 = b * sqrt(b);  // pow(b, 1.5)
        ke += ((rho2 + 2.0 * sigma2) * a1a2 + (x12a1 * x12a2)) / p15 / 16.0 / kPi;
        kef += (rho2 + 1.5 * sigma2) * a1a2 / p15 / 8.0 / kPi;
        if (HighOrder) {
          // ENSTROPHY / ENSTROPHYF pairs (interaction_utils.h:141-202) with
          // ENST / ENSTF (highorder.h:77-96); sqrt(b^7) as b^3 sqrt(b)
          const double sig2 = sigma * sigma;
          const double bb = rho2 + sig2;
          const double s7 = bb * bb * bb * sqrt(bb);
          const double factor = (k1Pi / 8.0) / s7;
          const double F1 = factor * (2.0 * rho2 * rho2 + 7.0 * sig2 * rho2 + 20.0 * sig2 * sig2);
          const double F2 = factor * (-6.0 * rho2 * rho2 - 27.0 * sig2 * rho2 - 21.0 * sig2 * sig2) / bb;
          en += F1 * a1a2 + F2 * x12a1 * x12a2;
          enf += (15.0 * k1Pi * sig2 * sig2 / 8.0) / s7 * a1a2;
        }
      }
    }
  }
  if (active && it.nself == 0) {
    // self terms (interaction_utils.h:249-267); items of a cross-wake block
    // (nself = 1) have none
    const double sg = radius[t];
    double a2 = 0.0;
    a2 += me.ax * me.ax;
    a2 += me.ay * me.ay;
    a2 += me.az * me.az;
    ke += k1Pi * a2 / sg / 8.0;
    kef += 3 * k1Pi * a2 / (sg * 16.0);
    if (HighOrder) {
      // ENST(sigma), ENSTF(sigma) (highorder.h:98-106)
      en += 5.0 * k1Pi / 2.0 / (sg * sg * sg) * a2;
      enf += 15.0 * k1Pi / 8.0 / (sg * sg * sg) * a2;
    }
  }
  double v[kDiagQuad] = {active ? h : 0.0, active ? ke : 0.0, active ? kef : 0.0,
                         active ? en : 0.0, active ? enf : 0.0};
  block_sum<kDiagQuad>(v, red);
  if (threadIdx.x == 0)
    for (int m = 0; m < kDiagQuad; ++m) out[kDiagQuad * blockIdx.x + m] = v[m];
}

}  // namespace

hipError_t launch_diag_linear(const Rec64* y, const WorkItem* items,
                              int nitems, double* out, hipStream_t st) {
  if (nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_diag_linear, dim3(nitems), dim3(kTB), 0, st, y, items, out);
  return hipGetLastError();
}

hipError_t launch_diag_pairs(const Rec64* y, const double* radius,
                             const WorkItem* items, int nitems, double* out,
                             bool high_order, hipStream_t st) {
  if (nitems <= 0) return hipSuccess;
  if (high_order)
    hipLaunchKernelGGL(k_diag_pairs<true>, dim3(nitems), dim3(kTB), 0, st, y, radius, items, out);
  else
    hipLaunchKernelGGL(k_diag_pairs<false>, dim3(nitems), dim3(kTB), 0, st, y, radius, items, out);
  return hipGetLastError();
}

}  // namespace pw
