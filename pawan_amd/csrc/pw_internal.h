// pw_internal.h -- device data layout and kernel launchers shared by the
// HIP translation units of libpawan_gpu.so.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pw {

// One particle of the device state: position, squared smoothing radius,
// vorticity strength and volume (64 B, one cache-line quarter).  The state
// of every stage of a step is an array of these; the pair kernels read it
// both as targets and, staged through LDS, as sources.
struct alignas(64) Rec64 {
  double x, y, z, s2;
  double ax, ay, az, vol;
};

// Rates of one particle: velocity u and vorticity rate w (48 B).
struct alignas(16) Rate64 {
  double ux, uy, uz, wx, wy, wz;
};

// One block of work for the pair kernels: targets [t0, t1) against
// sources [s0, s1).  t1 - t0 is at most the block's target count.  part < 0:
// the item covers all of its targets' sources and writes their rates; part
// >= 0: it covers one of several source ranges and writes a partial sum to
// slice `part` of the partial buffer.  (The diagnostics reuse the
// structure with {first, last, wake, -}.)
struct WorkItem {
  int32_t t0, t1, s0, s1;
  int32_t part, pad0, pad1, pad2;
};

// Arguments of a pair-kernel launch (passed by value).
struct PairArgs {
  const Rec64* state;
  const WorkItem* items;
  int nitems;
  int n;              // particles (stride of the partial buffer)
  Rate64* rates;
  Rate64* partial;    // parts x n, used by items with part >= 0
  int* queue;         // work counter of the persistent kernels
  double nu;
};

// How a right-hand side is launched (chosen by the autotuner).
struct PairConfig {
  int parts = 1;           // source ranges per target block
  bool persistent = false; // resident blocks pulling items from a queue
  bool waves8 = false;     // FP32C: register budget of 8 waves/SIMD
};

// Pair-kernel variants.
enum PairKind : int {
  kPairFP64Exact = 0,  // the reference's kernel formula for every pair
  kPairFP64 = 1,       // same, closed far-field form where exp/erf terms
                       // are below FP64 rounding (rho^2 > 90 sigma^2)
  kPairFP32C = 2,      // FP32 pair math, compensated FP32 accumulation
};

// Pair kernel of `kind` over args.items; for parts > 1 also the reduction
// of the partial sums into args.rates.  grid = resident blocks for the
// persistent form.
hipError_t launch_pair(int kind, const PairConfig& cfg, const PairArgs& args, int grid,
                       hipStream_t st);
// waves per SIMD the pair kernel is built for (for the persistent grid)
int pair_kernel_waves(int kind, const PairConfig& cfg);

// Stage arithmetic of the steppers (FP64, no contraction, the reference's
// operation order).  n = number of particles.
//   out = k * c + y                     (rk4.cpp:31-34 etc.)
hipError_t launch_stage(const Rec64* y, const Rate64* k, double c, Rec64* out,
                        int n, hipStream_t st);
//   y = y + ((k1 + k4) * c6 + (k2 + k3) * c3)      (rk4.cpp:58-66)
hipError_t launch_rk4_final(Rec64* y, const Rate64* k1, const Rate64* k2,
                            const Rate64* k3, const Rate64* k4, double c6,
                            double c3, int n, hipStream_t st);
//   k1 = k1 * dt ; out = (k2 * 2dt + y) - k1      (rk3.cpp:37-41)
hipError_t launch_rk3_mid(const Rec64* y, Rate64* k1, const Rate64* k2,
                          double dt, double two_dt, Rec64* out, int n,
                          hipStream_t st);
//   y = y + ((k1 * (1/6) + k3 * (dt/6)) + k2 * (dt*2/3))   (rk3.cpp:47-55)
hipError_t launch_rk3_final(Rec64* y, const Rate64* k1, const Rate64* k2,
                            const Rate64* k3, double c6, double c23, int n,
                            hipStream_t st);
//   y = y + k * dt                      (integration.cpp:130-133)
hipError_t launch_euler(Rec64* y, const Rate64* k, double dt, int n,
                        hipStream_t st);
// count of non-finite entries of the rates -> *flag (device int)
hipError_t launch_check_finite(const Rate64* k, int n, int* flag,
                               hipStream_t st);

// zero-fill of a small int buffer (a kernel, not hipMemsetAsync, so that
// the step sequence can be captured into a HIP graph)
hipError_t launch_fill_int(int* p, int n, int value, hipStream_t st);

// Diagnostics (FP64).  Linear sums: items[b] = {first, last, -, -} particle
// ranges inside one wake; out[11 * b + m] gets the block's sums of
// [a (3), x x a (3), x x (x x a) (3), |a| z (x^2+y^2), |a| (x^2+y^2)].
hipError_t launch_diag_linear(const Rec64* y, const WorkItem* items,
                              int nitems, double* out, hipStream_t st);
// Quadratic sums over the pair work items: out[3 * b + m] gets the block's
// total of [helicity, KE, KEF] rows (self terms included).
hipError_t launch_diag_pairs(const Rec64* y, const double* radius,
                             const WorkItem* items, int nitems, double* out,
                             hipStream_t st);

}  // namespace pw
