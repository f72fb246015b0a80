// pw_pair.hip -- tiled direct-sum pair kernels for gfx950.
//
// Direct sum: one work item = one block of 256 lanes (four wave64): up to
// 256 targets, one per lane, against a contiguous source range.  Sources are staged
// through LDS in tiles of 256 records, loaded with coalesced 16-byte global
// loads, and read back as wave-uniform (broadcast) LDS reads.  Every lane
// accumulates the rates of its own target in source order: the results are
// deterministic and need no atomics.
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
//
// The stretching q (a_t x a_s) + F [d . (a_t x a_s)] d is summed with the
// target's strength factored out, a_t x sum q a_s and d.(a_t x a_s) =
// a_t.(a_s x d) (the velocity's cross product), for every source that is
// beyond the Gaussian core of all the wave's targets; a source inside some
// lane's core adds a_t x a_s pair by pair, as the reference does.  Factoring
// those too gathers large, nearly parallel strengths into sum q a_s (thin
// rings), whose cross product with a_t then cancels several digits.
// Where rho^2 > kFar sigma^2 the Gaussian terms (exp(-rho_bar^2/2) and
// erfc(rho_bar/sqrt 2)) are below the arithmetic's rounding, and
// q = 1/(4 pi rho^3), F = -3 q / rho^2, Z = 0 exactly as far as the result
// can tell; only nearer pairs (about 1-2 % of them on the benchmark wakes)
// take the full kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>

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

// FP64 far field: beyond rho^2 > 90 sigma^2, exp(-rho_bar^2/2) and
// erfc(rho_bar/sqrt 2) are below 1e-17 of the algebraic term.
constexpr double kFar64 = 90.0;
// FP32 far field: the same terms are below 3e-8 beyond 44 sigma^2.
constexpr float kFar32 = 44.0f;
// Below rho_bar^2 < 4 the FP32 path evaluates q and F from positive
// series (no cancellation).
constexpr float kSeries32 = 4.0f;

// ---------------------------------------------------------------------
// FP64, reference form: the reference's formula, operation by operation,
// for every pair (PW_FP64_REF).
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

// Where an item's rates go: straight to the rates, or to its part's slice
// of the partial-sum buffer (source-split items, summed by k_reduce_parts).
__device__ __forceinline__ Rate64* out_rates(const PairArgs& a, const WorkItem& it) {
  return it.part < 0 ? a.rates : a.partial + static_cast<size_t>(it.part) * a.n;
}

__device__ __forceinline__ void item_fp64_ref(const PairArgs& pa, const WorkItem& it,
                                              Rec64* tile) {
  const Rec64* __restrict__ state = pa.state;
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
        pair64_exact(me, tile[l], pa.nu, a);
      }
    }
  }
  if (active) {
    Rate64 r;
    r.ux = a.ux; r.uy = a.uy; r.uz = a.uz;
    r.wx = a.wx; r.wy = a.wy; r.wz = a.wz;
    out_rates(pa, it)[t] = r;
  }
}

// ---------------------------------------------------------------------
// FP64 (PW_FP64): the reference's sums with the closed far-field form.
// ---------------------------------------------------------------------
struct T64 {
  double x, y, z, hs2, ax, ay, az, vol, nu2;
};

// 1/sqrt(x) for positive finite x (every pair the FP64 kernel evaluates has
// r^2 > 0): the hardware estimate y (relative error below 2^-22) and one
// third-order correction, y (1 + e/2 + 3e^2/8) with e = 1 - x y^2, which
// leaves O(e^3) < 2^-66 before rounding.  No special-value handling.
__device__ __forceinline__ double rsqrt_pos(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// Sums of the FP64 kernel.  A source beyond the Gaussian core of every
// lane's target adds its stretching with the target's strength factored
// out (d.(a_t x a_s) = a_t.(a_s x d), v = a_s x d shared with the velocity;
// sum q (a_t x a_s) = a_t x A); a source in some lane's core adds
// q (a_t x a_s) pair by pair into B, where A would gather large, nearly
// parallel strengths (thin rings) and a_t x A would cancel.
//   w = a_t x A + B,  A = sum q a_s,  B = sum F (a_t.v) d + per-pair terms + PSE.
struct AccF64 {
  double ux = 0, uy = 0, uz = 0;  // sum q v
  double ax = 0, ay = 0, az = 0;  // A
  double bx = 0, by = 0, bz = 0;  // B
};

__device__ __forceinline__ void pair64(const T64& t, const Rec64& s, AccF64& a) {
  const double dx = t.x - s.x, dy = t.y - s.y, dz = t.z - s.z;
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double sig2 = 0.5 * s.s2 + t.hs2;
  const double ir = rsqrt_pos(r2);
  const double ir2 = ir * ir;
  double q = kQ * (ir * ir2);
  double F = -3.0 * q * ir2;
  const bool core = r2 <= kFar64 * sig2;
  if (core) {
    const double isg = rsqrt_pos(sig2);
    const double isg3 = isg * isg * isg;
    const double rb = (r2 * ir) * isg;    // rho / sigma
    const double irb = (sig2 * isg) * ir;  // sigma / rho
    const double Z = exp(-0.5 * rb * rb) * isg3 * kK;
    const double phi = kQ * erf(kSqrtHalf * rb) * isg3;
    q = (phi * irb - Z) * (irb * irb);
    F = (Z - 3.0 * q) * ir2;
    const double c = t.nu2 * Z * (isg * isg);  // 2 nu Z / sigma^2
    a.bx += c * (t.vol * s.ax - s.vol * t.ax);
    a.by += c * (t.vol * s.ay - s.vol * t.ay);
    a.bz += c * (t.vol * s.az - s.vol * t.az);
  }
  const double vx = s.ay * dz - s.az * dy;
  const double vy = s.az * dx - s.ax * dz;
  const double vz = s.ax * dy - s.ay * dx;
  a.ux += q * vx;
  a.uy += q * vy;
  a.uz += q * vz;
  const double fr = F * (t.ax * vx + t.ay * vy + t.az * vz);
  if (__ballot(core)) {
    a.bx += q * (t.ay * s.az - t.az * s.ay) + fr * dx;
    a.by += q * (t.az * s.ax - t.ax * s.az) + fr * dy;
    a.bz += q * (t.ax * s.ay - t.ay * s.ax) + fr * dz;
  } else {
    a.ax += q * s.ax;
    a.ay += q * s.ay;
    a.az += q * s.az;
    a.bx += fr * dx;
    a.by += fr * dy;
    a.bz += fr * dz;
  }
}

// Wave reductions and block-uniform values.
__device__ __forceinline__ float wave_max(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ float wave_min(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ double wave_maxd(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ double wave_mind(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}
// a block-uniform value held in scalar registers
__device__ __forceinline__ float uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uniform(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane(static_cast<int>(b & 0xffffffff));
  const int hi = __builtin_amdgcn_readfirstlane(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Separated source ranges of the FP64 kernel.  Positions relative to the
// block's targets' centre c_T (t' = x_t - c_T, r_s = x_s - c_T, FP64), with
// m_s = a_s x r_s staged with each source and c_t = t' x a_t per target:
//   sum q' (a_s x d) = A x t' - M,   a_t.(a_s x d) = a_s.c_t - a_t.m_s,
// A = sum q' a_s, M = sum q' m_s, B = sum f' (a_t.(a_s x d)) d, with
// q' = 1/rho^3 and f' = q'/rho^2 (the singular law; the range is beyond
// 90 sigma^2 of every pair, where the Gaussian terms are below FP64
// rounding), and u = (A x t' - M)/4pi, w = (a_t x A - 3 B)/4pi.  The
// subtractions cancel at most the distances to c_T over the pair
// distance, which the separation test bounds by kSepFP64: 2^-53 x 16 of
// each term.
constexpr double kSepFP64 = 16.0;

struct TileF64 {
  Rec64 rec;             // the state record (the core-capable pairs)
  double rx, ry, rz;     // x - c_T
  double mx, my, mz;     // a x (x - c_T)
};

struct P64S {
  double ax = 0, ay = 0, az = 0;  // A
  double mx = 0, my = 0, mz = 0;  // M
  double bx = 0, by = 0, bz = 0;  // B
};

// one separated pair: (tx, ty, tz) = t', (cx, cy, cz) = t' x a_t; the
// source's r, a and m = a x r
__device__ __forceinline__ void pair64s(double tx, double ty, double tz, double tax, double tay, double taz,
                                        double cx, double cy, double cz, double srx, double sry, double srz,
                                        double sax, double say, double saz, double smx, double smy, double smz,
                                        P64S& p) {
  const double dx = tx - srx, dy = ty - sry, dz = tz - srz;
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double ir = rsqrt_pos(r2);
  const double ir2 = ir * ir;
  const double q = ir * ir2;
  const double f = q * ir2;
  p.ax += q * sax;
  p.ay += q * say;
  p.az += q * saz;
  p.mx += q * smx;
  p.my += q * smy;
  p.mz += q * smz;
  const double h = (sax * cx + say * cy + saz * cz) - (tax * smx + tay * smy + taz * smz);
  const double g = f * h;
  p.bx += g * dx;
  p.by += g * dy;
  p.bz += g * dz;
}

__device__ __forceinline__ void item_fp64(const PairArgs& pa, const WorkItem& it, TileF64* tile) {
  __shared__ double s_tb[kTB / 64][6];  // each wave's targets' box
  __shared__ double s_b[kTB / 64][7];   // each staged chunk's box (relative to c_T) and largest sigma^2/2
  const Rec64* __restrict__ state = pa.state;
  const int lane = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
  const int t = it.t0 + static_cast<int>(threadIdx.x);
  const bool active = t < it.t1;
  const Rec64 m = state[active ? t : it.t0];
  T64 me;
  me.x = m.x; me.y = m.y; me.z = m.z; me.hs2 = 0.5 * m.s2;
  me.ax = m.ax; me.ay = m.ay; me.az = m.az; me.vol = m.vol;
  me.nu2 = 2.0 * pa.nu;
  // the block's and this wave's targets (inactive lanes repeat the first)
  {
    const double v[6] = {wave_mind(m.x), wave_mind(m.y), wave_mind(m.z),
                         wave_maxd(m.x), wave_maxd(m.y), wave_maxd(m.z)};
    __syncthreads();
    if (lane == 0)
      for (int k = 0; k < 6; ++k) s_tb[wave][k] = v[k];
  }
  __syncthreads();
  double cT[3], rT2 = 0.0, cW[3], rW2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double lo = fmin(fmin(s_tb[0][k], s_tb[1][k]), fmin(s_tb[2][k], s_tb[3][k]));
    const double hi = fmax(fmax(s_tb[0][3 + k], s_tb[1][3 + k]), fmax(s_tb[2][3 + k], s_tb[3][3 + k]));
    cT[k] = uniform(0.5 * (lo + hi));
    rT2 += 0.25 * (hi - lo) * (hi - lo);
    const double wl = s_tb[wave][k], wh = s_tb[wave][3 + k];
    cW[k] = uniform(0.5 * (wl + wh) - cT[k]);
    rW2 += 0.25 * (wh - wl) * (wh - wl);
  }
  const double rT = uniform(sqrt(rT2) * 1.000001), rW = uniform(sqrt(rW2) * 1.000001);
  const double h2T = uniform(wave_maxd(me.hs2));
  const double tx = m.x - cT[0], ty = m.y - cT[1], tz = m.z - cT[2];
  const double ctx = ty * me.az - tz * me.ay, cty = tz * me.ax - tx * me.az, ctz = tx * me.ay - ty * me.ax;
  AccF64 a;
  P64S sp;
  for (int base = it.s0; base < it.s1; base += kTB) {
    const int cnt = min(kTB, it.s1 - base);
    __syncthreads();
    double bl[3] = {INFINITY, INFINITY, INFINITY}, bh[3] = {-INFINITY, -INFINITY, -INFINITY}, h2 = 0.0;
    if (static_cast<int>(threadIdx.x) < cnt) {
      const Rec64 r = state[base + threadIdx.x];
      TileF64 f;
      f.rec = r;
      f.rx = r.x - cT[0];
      f.ry = r.y - cT[1];
      f.rz = r.z - cT[2];
      f.mx = r.ay * f.rz - r.az * f.ry;
      f.my = r.az * f.rx - r.ax * f.rz;
      f.mz = r.ax * f.ry - r.ay * f.rx;
      tile[threadIdx.x] = f;
      bl[0] = bh[0] = f.rx;
      bl[1] = bh[1] = f.ry;
      bl[2] = bh[2] = f.rz;
      h2 = 0.5 * r.s2;
    }
    {
      const double v[7] = {wave_mind(bl[0]), wave_mind(bl[1]), wave_mind(bl[2]), wave_maxd(bh[0]),
                           wave_maxd(bh[1]), wave_maxd(bh[2]), wave_maxd(h2)};
      if (lane == 0)
        for (int k = 0; k < 7; ++k) s_b[wave][k] = v[k];
    }
    __syncthreads();
    if (!active) continue;
    const bool diag = base == it.t0;
    const int self = static_cast<int>(threadIdx.x);
    for (int c0 = 0; c0 < cnt; c0 += 64) {
      const int c1 = min(cnt, c0 + 64);
      const int cw = c0 / 64;
      // the chunk's sphere (relative to c_T) against this wave's
      double cs2 = 0.0, rs2 = 0.0, d2 = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double l = s_b[cw][k], h = s_b[cw][3 + k];
        const double c = 0.5 * (l + h), e = 0.5 * (h - l);
        cs2 += c * c;
        rs2 += e * e;
        d2 += (c - cW[k]) * (c - cW[k]);
      }
      const double rS = uniform(sqrt(rs2) * 1.000001);
      const double gap = uniform(sqrt(d2) * 0.999999 - rW - rS);
      const bool sep = gap > 0.0 && gap * gap > kFar64 * (h2T + uniform(s_b[cw][6])) * 1.0001 &&
                       rT + uniform(sqrt(cs2)) + rS <= kSepFP64 * gap;
      if (sep) {
        for (int l = c0; l < c1; ++l) {
          const TileF64& f = tile[l];
          pair64s(tx, ty, tz, me.ax, me.ay, me.az, ctx, cty, ctz, f.rx, f.ry, f.rz, f.rec.ax, f.rec.ay, f.rec.az,
                  f.mx, f.my, f.mz, sp);
        }
        continue;
      }
      const bool own = diag && cw == wave;  // the chunk holding this wave's own targets
      for (int l = c0; l < c1; ++l)
        if (!(own && l == self)) pair64(me, tile[l].rec, a);  // the self pair is never evaluated
    }
  }
  if (active) {
    // separated ranges: u = (A x t' - M)/4pi, w = (a_t x A - 3 B)/4pi
    const double ux = (sp.ay * tz - sp.az * ty) - sp.mx;
    const double uy = (sp.az * tx - sp.ax * tz) - sp.my;
    const double uz = (sp.ax * ty - sp.ay * tx) - sp.mz;
    const double wx = (me.ay * sp.az - me.az * sp.ay) - 3.0 * sp.bx;
    const double wy = (me.az * sp.ax - me.ax * sp.az) - 3.0 * sp.by;
    const double wz = (me.ax * sp.ay - me.ay * sp.ax) - 3.0 * sp.bz;
    Rate64 r;
    r.ux = a.ux + kQ * ux;
    r.uy = a.uy + kQ * uy;
    r.uz = a.uz + kQ * uz;
    r.wx = ((me.ay * a.az - me.az * a.ay) + a.bx) + kQ * wx;
    r.wy = ((me.az * a.ax - me.ax * a.az) + a.by) + kQ * wy;
    r.wz = ((me.ax * a.ay - me.ay * a.ax) + a.bz) + kQ * wz;
    out_rates(pa, it)[t] = r;
  }
}

// ---------------------------------------------------------------------
// FP64 with the high-order algebraic kernel (src/interaction/highorder.h,
// the reference's compile-time alternative to the Gaussian).  With
// b = rho^2 + sigma^2 its KERNEL (highorder.h:56-67) is, in closed form,
//   q   = (1/4pi) (rho^2 + 5/2 sigma^2) b^-5/2,
//   F   = (Z - 3q)/rho^2 = -(3/8pi) (2 rho^2 + 7 sigma^2) b^-7/2,
//   eta = (105/8pi) sigma^6 b^-9/2,
// which the reference evaluates through rho_bar = rho/sigma and pow(); here
// one reciprocal square root of b gives all three, F without the
// cancellation of Z - 3q at small rho.  The kernel has no far cutoff (it
// decays algebraically), so every pair evaluates it.  The sums take the
// FP64 Gaussian kernel's per-(wave, source) choice: a source within
// rho^2 <= 90 sigma^2 of some lane's target adds its stretching
// q (a_t x a_s) and PSE c (V_t a_s - V_s a_t) pair by pair; any other source
// adds them factored, A += q a_s, C += c a_s, D += c V_s, and the item ends
// with w = a_t x A + B + V_t C - D a_t.
// ---------------------------------------------------------------------
constexpr double kHoF = -0x1.e8ec8a4aeacc4p-4;   // -3/(8 pi)
constexpr double kHoEta = 0x1.0b615ba0f867bp+2;  // 105/(8 pi)

struct AccHO {
  double ux = 0, uy = 0, uz = 0;  // sum q v
  double ax = 0, ay = 0, az = 0;  // A = sum q a_s (factored sources)
  double bx = 0, by = 0, bz = 0;  // B
  double cx = 0, cy = 0, cz = 0;  // C = sum c a_s (factored sources)
  double d = 0;                   // D = sum c V_s (factored sources)
};

__device__ __forceinline__ void pair64_ho(const T64& t, const Rec64& s, AccHO& a) {
  const double dx = t.x - s.x, dy = t.y - s.y, dz = t.z - s.z;
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double sig2 = 0.5 * s.s2 + t.hs2;
  const double ib = rsqrt_pos(r2 + sig2);
  const double ib2 = ib * ib;
  const double ib5 = ib2 * ib2 * ib;
  const double ib7 = ib5 * ib2;
  const double q = kQ * fma(2.5, sig2, r2) * ib5;
  const double F = kHoF * fma(7.0, sig2, 2.0 * r2) * ib7;
  const double c = t.nu2 * kHoEta * (sig2 * sig2) * (ib7 * ib2);  // 2 nu eta / sigma^2
  const double vx = s.ay * dz - s.az * dy;
  const double vy = s.az * dx - s.ax * dz;
  const double vz = s.ax * dy - s.ay * dx;
  a.ux += q * vx;
  a.uy += q * vy;
  a.uz += q * vz;
  const double fr = F * (t.ax * vx + t.ay * vy + t.az * vz);
  if (__ballot(r2 <= kFar64 * sig2)) {
    a.bx += q * (t.ay * s.az - t.az * s.ay) + fr * dx + c * (t.vol * s.ax - s.vol * t.ax);
    a.by += q * (t.az * s.ax - t.ax * s.az) + fr * dy + c * (t.vol * s.ay - s.vol * t.ay);
    a.bz += q * (t.ax * s.ay - t.ay * s.ax) + fr * dz + c * (t.vol * s.az - s.vol * t.az);
  } else {
    a.ax += q * s.ax;
    a.ay += q * s.ay;
    a.az += q * s.az;
    a.bx += fr * dx;
    a.by += fr * dy;
    a.bz += fr * dz;
    a.cx += c * s.ax;
    a.cy += c * s.ay;
    a.cz += c * s.az;
    a.d += c * s.vol;
  }
}

__device__ __forceinline__ void item_fp64_ho(const PairArgs& pa, const WorkItem& it, Rec64* tile) {
  const Rec64* __restrict__ state = pa.state;
  const int t = it.t0 + static_cast<int>(threadIdx.x);
  const bool active = t < it.t1;
  const Rec64 m = state[active ? t : it.t0];
  T64 me;
  me.x = m.x; me.y = m.y; me.z = m.z; me.hs2 = 0.5 * m.s2;
  me.ax = m.ax; me.ay = m.ay; me.az = m.az; me.vol = m.vol;
  me.nu2 = 2.0 * pa.nu;
  AccHO a;
  for (int base = it.s0; base < it.s1; base += kTB) {
    const int cnt = min(kTB, it.s1 - base);
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < cnt) tile[threadIdx.x] = state[base + threadIdx.x];
    __syncthreads();
    if (!active) continue;
    if (base == it.t0) {
      const int self = static_cast<int>(threadIdx.x);
      for (int l = 0; l < cnt; ++l)
        if (l != self) pair64_ho(me, tile[l], a);
    } else {
      for (int l = 0; l < cnt; ++l) pair64_ho(me, tile[l], a);
    }
  }
  if (active) {
    Rate64 r;
    r.ux = a.ux; r.uy = a.uy; r.uz = a.uz;
    r.wx = ((me.ay * a.az - me.az * a.ay) + a.bx) + (me.vol * a.cx - a.d * me.ax);
    r.wy = ((me.az * a.ax - me.ax * a.az) + a.by) + (me.vol * a.cy - a.d * me.ay);
    r.wz = ((me.ax * a.ay - me.ay * a.ax) + a.bz) + (me.vol * a.cz - a.d * me.az);
    out_rates(pa, it)[t] = r;
  }
}

// ---------------------------------------------------------------------
// Point-evaluation velocity (the reference's getVi,
// src/interaction/interaction.cpp:453-476): the velocity at m arbitrary
// points from every particle, with each particle's own radius as sigma,
// in FP64 with the closed far form beyond 90 sigma^2.  One lane per point,
// sources through LDS in tiles of 256 as in the pair kernels.  A point that
// coincides with a particle gets no contribution from it (a_s x d = 0).
// HighOrder: the algebraic kernel's q (highorder.h:31-39) instead.
// ---------------------------------------------------------------------
template <bool HighOrder>
__global__ __launch_bounds__(kTB) void k_point_velocity(const Rec64* __restrict__ state, int n,
                                                        const double* __restrict__ pts, int m,
                                                        double* __restrict__ u) {
  __shared__ Rec64 tile[kTB];
  const int k = blockIdx.x * kTB + static_cast<int>(threadIdx.x);
  const bool active = k < m;
  const double px = active ? pts[3 * k] : 0.0, py = active ? pts[3 * k + 1] : 0.0,
               pz = active ? pts[3 * k + 2] : 0.0;
  double ux = 0.0, uy = 0.0, uz = 0.0;
  for (int base = 0; base < n; base += kTB) {
    const int cnt = min(kTB, n - base);
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < cnt) tile[threadIdx.x] = state[base + threadIdx.x];
    __syncthreads();
    if (!active) continue;
    for (int l = 0; l < cnt; ++l) {
      const Rec64& s = tile[l];
      const double dx = px - s.x, dy = py - s.y, dz = pz - s.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 == 0.0) continue;
      double q;
      if (HighOrder) {
        const double ib = rsqrt_pos(r2 + s.s2);
        const double ib2 = ib * ib;
        q = kQ * fma(2.5, s.s2, r2) * (ib2 * ib2 * ib);
      } else {
      const double ir = rsqrt_pos(r2);
      const double ir2 = ir * ir;
      q = kQ * (ir * ir2);
      if (r2 <= kFar64 * s.s2) {
        const double isg = rsqrt_pos(s.s2);
        const double isg3 = isg * isg * isg;
        const double rb = (r2 * ir) * isg;    // rho / sigma
        const double irb = (s.s2 * isg) * ir;  // sigma / rho
        const double Z = exp(-0.5 * rb * rb) * isg3 * kK;
        const double phi = kQ * erf(kSqrtHalf * rb) * isg3;
        q = (phi * irb - Z) * (irb * irb);
      }
      }
      ux += q * (s.ay * dz - s.az * dy);
      uy += q * (s.az * dx - s.ax * dz);
      uz += q * (s.ax * dy - s.ay * dx);
    }
  }
  if (active) {
    u[3 * k] = ux;
    u[3 * k + 1] = uy;
    u[3 * k + 2] = uz;
  }
}

// ---------------------------------------------------------------------
// FP32 pair math, compensated FP32 accumulation, FP64 state (PW_FP32C).
// ---------------------------------------------------------------------
// Source record in LDS: FP64 position (the displacement is formed in FP64
// and rounded once: a ring's velocity is a small residual of large
// opposing contributions, and rounding positions first would cost it),
// FP32 half squared radius, strength and volume.  48 B, three 16-byte
// broadcast reads per pair.
struct alignas(16) SrcT {
  double x, y, z;
  float hs2, ax, ay, az, vol, pad;
};

// Partial sums, with 1/(4 pi) factored out of q and F:
// q' = 4 pi q, f' = -(4 pi / 3) F  (far field: q' = rho^-3, f' = rho^-5).
// The stretching of a pair, q' (a_t x a_s) - 3 f' [d.(a_t x a_s)] d, is
// summed with the target's strength factored out: d.(a_t x a_s) =
// a_t.(a_s x d) reuses the velocity's cross product v = a_s x d, and
// sum q' (a_t x a_s) = a_t x sum q' a_s, so a pair costs no cross product
// of its own; fold() forms w = a_t x A - 3 B once per chunk.
struct P32 {
  float ux = 0, uy = 0, uz = 0;  // U = sum q' v
  float ax = 0, ay = 0, az = 0;  // A = sum q' a_s
  float bx = 0, by = 0, bz = 0;  // B = sum f' (a_t.v) d
};
// PSE sums collect only the near pairs (about 1 % of them) and are summed
// directly.
struct PSE32 {
  float x = 0, y = 0, z = 0;  // sum c (V_t a_s - V_s a_t)
};
// the same sums with the target's factors taken out: (x, y, z) = sum c a_s,
// w = sum c V_s, so that the sum is V_t (x, y, z) - w a_t
struct PSE32F {
  float x = 0, y = 0, z = 0, w = 0;
};
constexpr int kAcc = 6;
constexpr int kChunk = 64;  // pairs summed in plain FP32 before compensation

struct T32 {
  double x, y, z;
  float hs2, ax, ay, az, vol, nu2;
};

constexpr float kFourPiK = 0x1.988454p-1f;   // 4 pi (2 pi)^-1.5 = sqrt(2/pi)
constexpr float kFourPi = 0x1.921fb6p+3f;    // 4 pi
constexpr float kFourPi3 = 0x1.0c1524p+2f;   // 4 pi / 3
constexpr float kKf = 0x1.0411e8p-4f;        // (2 pi)^-1.5
constexpr float kQf = 0x1.45f306p-4f;        // 1/(4 pi)

// Near field of the FP32 kernel (rho^2 <= 44 sigma^2): q', f' and the PSE
// weight c, with
//   s = rho_bar^2, e = exp(-s/2), B(s) = sum_k s^k / (2k+5)!!,
//   4 pi q = sqrt(2/pi) e (1/3 + s B) / sigma^3,
//   F      = -3 (2 pi)^-1.5 e B / sigma^5
// for s < 4 (the incomplete-gamma series of erf, A&S 6.5.29: every term
// positive, so no cancellation as rho -> 0), and the closed erf form for
// 4 <= s <= 44 where its cancellation is mild.
// e = exp(-s/2) as 2^(-s log2(e) / 2) on the hardware exp2: the rounding of
// the argument (|arg| <= 32) costs at most 2^-24 |arg| ln 2 relative, and
// at large s the Gaussian terms are a vanishing part of q.
constexpr float kMHalfLog2e = -0x1.715476p-1f;  // -log2(e) / 2
// erfc(x) = exp(-x^2) R(w) for x = rho / (sqrt(2) sigma), w = sigma / rho:
// R = exp(x^2) erfc(x) (the scaled complementary error function) as a
// degree-6 polynomial in w on 4 <= rho^2/sigma^2 <= 44, a Chebyshev least-
// squares fit (fit_erfcx.py beside this file); largest relative error of
// the FP32 Horner evaluation 3.8e-7, i.e. below 2e-8 absolute in erf.  The
// exponential is the e above, since x^2 = s/2.
constexpr float kErfcx[7] = {0x1.c20ea2p-13f, 0x1.9559b4p-1f, 0x1.354a9p-4f, -0x1.4d4cdcp+0f,
                             0x1.efececp+0f, -0x1.65f268p+0f, 0x1.b22f5p-2f};

__device__ __forceinline__ void near32(float r2, float sig2, float nu2, float& qp,
                                       float& fp, float& c) {
  const float isg = __builtin_amdgcn_rsqf(sig2);
  const float isg2 = isg * isg;
  const float isg3 = isg2 * isg;
  const float sb = r2 * isg2;
  const float e = __builtin_amdgcn_exp2f(kMHalfLog2e * sb);
  if (sb < kSeries32) {
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
    const float ke = kFourPiK * e * isg3;
    qp = ke * A;
    fp = ke * B * isg2;  // -(4 pi/3) F = 4 pi K e B / sigma^5
  } else {
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float irb = sig2 * isg * ir;  // w = sigma / rho
    const float Z = kKf * e * isg3;
    float R = kErfcx[6];
#pragma unroll
    for (int k = 5; k >= 0; --k) R = fmaf(R, irb, kErfcx[k]);
    const float phi = kQf * fmaf(-e, R, 1.0f) * isg3;  // erf(x) / (4 pi sigma^3)
    const float q = (phi * irb - Z) * (irb * irb);
    const float F = (Z - 3.0f * q) * (ir * ir);
    qp = kFourPi * q;
    fp = -kFourPi3 * F;
  }
  c = nu2 * kKf * e * isg3 * isg2;  // 2 nu Z / sigma^2
}

// The Gaussian core without branches, for the tree path's near field:
//   G(s) = exp(-s/2) B(s),  q' = sqrt(2/pi) (e/3 + s G) / sigma^3,
//   f' = sqrt(2/pi) G / sigma^5,
// the same series as near32 with the exponential factored, so one function
// serves every lane whatever its s (near32 runs both of its branches when a
// wave's pairs straddle s = 4).  G comes from a table of 64 pieces of degree
// 4 in t = s - s_c on 0 <= s <= 44 (rows (c0 .. c4, 0, s_c, 0); a Chebyshev
// fit, fit_core_table.py beside this file; largest relative error of this
// FP32 evaluation 1.2e-7), staged in LDS and read per lane.
constexpr int kCorePieces = 64;
constexpr float kCoreScale = 0x1.74539p+0f;  // kCorePieces / (44 * 1.0001) in FP32
__constant__ float kCoreTable[64][8] = {
    {0x1.e35bd2p-5f, -0x1.557002p-6f, 0x1.07b0e8p-8f, -0x1.1ea542p-11f, 0x1.e3402ap-15f, 0.0f, 0x1.600902p-2f, 0.0f},
    {0x1.7c34aap-5f, -0x1.0654f4p-6f, 0x1.8f30cap-9f, -0x1.ad7d32p-12f, 0x1.675396p-15f, 0.0f, 0x1.0806c2p+0f, 0.0f},
    {0x1.2cccb2p-5f, -0x1.94bdf8p-7f, 0x1.2f0d26p-9f, -0x1.4279ep-12f, 0x1.0ba584p-15f, 0.0f, 0x1.b80b44p+0f, 0.0f},
    {0x1.ded68ap-6f, -0x1.399146p-7f, 0x1.cd91dcp-10f, -0x1.e56572p-13f, 0x1.8f72d4p-16f, 0.0f, 0x1.3407e2p+1f, 0.0f},
    {0x1.7f870cp-6f, -0x1.e80d76p-8f, 0x1.60a9bap-10f, -0x1.6e393cp-13f, 0x1.2aa7b8p-16f, 0.0f, 0x1.8c0a24p+1f, 0.0f},
    {0x1.353016p-6f, -0x1.7d9c8cp-8f, 0x1.0e6414p-10f, -0x1.15095ap-13f, 0x1.bf807p-17f, 0.0f, 0x1.e40c64p+1f, 0.0f},
    {0x1.f5d95p-7f, -0x1.2bdc3ap-8f, 0x1.a024cep-11f, -0x1.a44d18p-14f, 0x1.4ffcap-17f, 0.0f, 0x1.1e0752p+2f, 0.0f},
    {0x1.9a0f8p-7f, -0x1.d9ac1p-9f, 0x1.4177p-11f, -0x1.3fc1eap-14f, 0x1.f9aa08p-18f, 0.0f, 0x1.4a0872p+2f, 0.0f},
    {0x1.51611cp-7f, -0x1.781d8p-9f, 0x1.f2a82ep-12f, -0x1.e806a2p-15f, 0x1.7d6c0cp-18f, 0.0f, 0x1.760994p+2f, 0.0f},
    {0x1.1783eep-7f, -0x1.2c4c6ep-9f, 0x1.846394p-12f, -0x1.759f86p-15f, 0x1.206d6ap-18f, 0.0f, 0x1.a20ab4p+2f, 0.0f},
    {0x1.d26158p-8f, -0x1.e23f76p-10f, 0x1.2fd524p-12f, -0x1.1f0176p-15f, 0x1.b55c1ep-19f, 0.0f, 0x1.ce0bd4p+2f, 0.0f},
    {0x1.87c7b2p-8f, -0x1.8576acp-10f, 0x1.dd8968p-13f, -0x1.ba7f4ep-16f, 0x1.4c837p-19f, 0.0f, 0x1.fa0cf4p+2f, 0.0f},
    {0x1.4b5cbep-8f, -0x1.3c6274p-10f, 0x1.790b48p-13f, -0x1.565fbep-16f, 0x1.fb10e8p-20f, 0.0f, 0x1.13070ap+3f, 0.0f},
    {0x1.1a23f4p-8f, -0x1.028ceap-10f, 0x1.2b2546p-13f, -0x1.09ec68p-16f, 0x1.83cbaap-20f, 0.0f, 0x1.29079ap+3f, 0.0f},
    {0x1.e3998p-9f, -0x1.a91c6ep-11f, 0x1.dd0cc4p-14f, -0x1.9ebc6ep-17f, 0x1.298474p-20f, 0.0f, 0x1.3f082ap+3f, 0.0f},
    {0x1.a11428p-9f, -0x1.5f955ep-11f, 0x1.7e5112p-14f, -0x1.44bf98p-17f, 0x1.ca043cp-21f, 0.0f, 0x1.5508bap+3f, 0.0f},
    {0x1.69e77ap-9f, -0x1.2482aap-11f, 0x1.33fc3cp-14f, -0x1.febd4ep-18f, 0x1.61c222p-21f, 0.0f, 0x1.6b094ap+3f, 0.0f},
    {0x1.3bde48p-9f, -0x1.e99abep-12f, 0x1.f2d1c4p-15f, -0x1.9362fcp-18f, 0x1.123466p-21f, 0.0f, 0x1.8109dcp+3f, 0.0f},
    {0x1.153b5p-9f, -0x1.9c236p-12f, 0x1.96174p-15f, -0x1.4007b8p-18f, 0x1.aaa61p-22f, 0.0f, 0x1.970a6cp+3f, 0.0f},
    {0x1.e93c4ep-10f, -0x1.5ce924p-12f, 0x1.4c5c2p-15f, -0x1.fe20aap-19f, 0x1.4d2f48p-22f, 0.0f, 0x1.ad0afcp+3f, 0.0f},
    {0x1.b1dd48p-10f, -0x1.2907acp-12f, 0x1.117682p-15f, -0x1.98783cp-19f, 0x1.0536f6p-22f, 0.0f, 0x1.c30b8cp+3f, 0.0f},
    {0x1.829a0ap-10f, -0x1.fc7794p-13f, 0x1.c46378p-16f, -0x1.489e34p-19f, 0x1.9b3a7ap-23f, 0.0f, 0x1.d90c1cp+3f, 0.0f},
    {0x1.5a0b7cp-10f, -0x1.b57d3cp-13f, 0x1.7827bap-16f, -0x1.09a3c6p-19f, 0x1.45077ap-23f, 0.0f, 0x1.ef0cacp+3f, 0.0f},
    {0x1.371018p-10f, -0x1.7a52dcp-13f, 0x1.3a6318p-16f, -0x1.af869ap-20f, 0x1.01fa28p-23f, 0.0f, 0x1.02869ep+4f, 0.0f},
    {0x1.18bd54p-10f, -0x1.48c058p-13f, 0x1.08193ep-16f, -0x1.603078p-20f, 0x1.9b4378p-24f, 0.0f, 0x1.0d86e6p+4f, 0.0f},
    {0x1.fca886p-11f, -0x1.1f01e8p-13f, 0x1.bdeb1ep-17f, -0x1.20d1e8p-20f, 0x1.493bcp-24f, 0.0f, 0x1.18872ep+4f, 0.0f},
    {0x1.ce7176p-11f, -0x1.f75d7cp-14f, 0x1.7a4942p-17f, -0x1.dbf7c6p-21f, 0x1.08b764p-24f, 0.0f, 0x1.238776p+4f, 0.0f},
    {0x1.a5d4c6p-11f, -0x1.bb4ae4p-14f, 0x1.426dbcp-17f, -0x1.8a0c18p-21f, 0x1.ab8e5ep-25f, 0.0f, 0x1.2e87bep+4f, 0.0f},
    {0x1.81fe82p-11f, -0x1.87f9b4p-14f, 0x1.141434p-17f, -0x1.47c004p-21f, 0x1.5ace54p-25f, 0.0f, 0x1.398806p+4f, 0.0f},
    {0x1.623eccp-11f, -0x1.5bf098p-14f, 0x1.dae166p-18f, -0x1.11dc2cp-21f, 0x1.1a8bfap-25f, 0.0f, 0x1.44884ep+4f, 0.0f},
    {0x1.46031p-11f, -0x1.35fcep-14f, 0x1.9a2a86p-18f, -0x1.cbb984p-22f, 0x1.ce690ap-26f, 0.0f, 0x1.4f8896p+4f, 0.0f},
    {0x1.2cd092p-11f, -0x1.152434p-14f, 0x1.63bb08p-18f, -0x1.839142p-22f, 0x1.7c093cp-26f, 0.0f, 0x1.5a88dep+4f, 0.0f},
    {0x1.164016p-11f, -0x1.f1329ep-15f, 0x1.35bde6p-18f, -0x1.4824p-22f, 0x1.39af52p-26f, 0.0f, 0x1.658926p+4f, 0.0f},
    {0x1.01fa6ap-11f, -0x1.bf6642p-15f, 0x1.0eb95ep-18f, -0x1.16fcfp-22f, 0x1.04057ep-26f, 0.0f, 0x1.70897p+4f, 0.0f},
    {0x1.df6b38p-12f, -0x1.93cbd6p-15f, 0x1.daf8p-19f, -0x1.dc52b2p-23f, 0x1.b0e2c4p-27f, 0.0f, 0x1.7b89b8p+4f, 0.0f},
    {0x1.be6548p-12f, -0x1.6d7b1ap-15f, 0x1.a21b7cp-19f, -0x1.98381p-23f, 0x1.69d1eap-27f, 0.0f, 0x1.868ap+4f, 0.0f},
    {0x1.a0774p-12f, -0x1.4bb264p-15f, 0x1.7148bcp-19f, -0x1.5f2f1ap-23f, 0x1.2fa402p-27f, 0.0f, 0x1.918a48p+4f, 0.0f},
    {0x1.854498p-12f, -0x1.2dcf9p-15f, 0x1.4734bp-19f, -0x1.2f3a74p-23f, 0x1.ffa41ep-28f, 0.0f, 0x1.9c8a9p+4f, 0.0f},
    {0x1.6c7d96p-12f, -0x1.134a46p-15f, 0x1.22cf44p-19f, -0x1.06bfeap-23f, 0x1.b0b834p-28f, 0.0f, 0x1.a78ad8p+4f, 0.0f},
    {0x1.55dd4p-12f, -0x1.f75eeap-16f, 0x1.0337f6p-19f, -0x1.c8e7b6p-24f, 0x1.6f572ap-28f, 0.0f, 0x1.b28b2p+4f, 0.0f},
    {0x1.4127bep-12f, -0x1.cd3b44p-16f, 0x1.cf69aap-20f, -0x1.8e9254p-24f, 0x1.38f8fep-28f, 0.0f, 0x1.bd8b68p+4f, 0.0f},
    {0x1.2e28fp-12f, -0x1.a783ccp-16f, 0x1.9f568p-20f, -0x1.5cc942p-24f, 0x1.0b977p-28f, 0.0f, 0x1.c88bbp+4f, 0.0f},
    {0x1.1cb34ep-12f, -0x1.85ab08p-16f, 0x1.753456p-20f, -0x1.32260ep-24f, 0x1.cb250cp-29f, 0.0f, 0x1.d38bf8p+4f, 0.0f},
    {0x1.0c9efp-12f, -0x1.6737dp-16f, 0x1.502af6p-20f, -0x1.0d828ap-24f, 0x1.8b3664p-29f, 0.0f, 0x1.de8c4p+4f, 0.0f},
    {0x1.fb9186p-13f, -0x1.4bc208p-16f, 0x1.2f849cp-20f, -0x1.dbd8d6p-25f, 0x1.55452ep-29f, 0.0f, 0x1.e98c88p+4f, 0.0f},
    {0x1.e023b8p-13f, -0x1.32efeep-16f, 0x1.12a7fp-20f, -0x1.a536dp-25f, 0x1.2799f4p-29f, 0.0f, 0x1.f48cdp+4f, 0.0f},
    {0x1.c6bddp-13f, -0x1.1c73eap-16f, 0x1.f2263ap-21f, -0x1.75d2bp-25f, 0x1.00ceeep-29f, 0.0f, 0x1.ff8d18p+4f, 0.0f},
    {0x1.af2f5ap-13f, -0x1.080aaap-16f, 0x1.c4afcp-21f, -0x1.4c97d4p-25f, 0x1.bf7fap-30f, 0.0f, 0x1.0546bp+5f, 0.0f},
    {0x1.994d4cp-13f, -0x1.eaf33ep-17f, 0x1.9c3088p-21f, -0x1.289f2ap-25f, 0x1.86f9f8p-30f, 0.0f, 0x1.0ac6d4p+5f, 0.0f},
    {0x1.84f14ep-13f, -0x1.c91b62p-17f, 0x1.7807fp-21f, -0x1.092704p-25f, 0x1.5682acp-30f, 0.0f, 0x1.1046f8p+5f, 0.0f},
    {0x1.71f928p-13f, -0x1.aa345ep-17f, 0x1.57ac3cp-21f, -0x1.db191ap-26f, 0x1.2cd45ep-30f, 0.0f, 0x1.15c71cp+5f, 0.0f},
    {0x1.604636p-13f, -0x1.8defcep-17f, 0x1.3aa6fcp-21f, -0x1.aa8ccep-26f, 0x1.08e19cp-30f, 0.0f, 0x1.1b474p+5f, 0.0f},
    {0x1.4fbd02p-13f, -0x1.7408cp-17f, 0x1.20920cp-21f, -0x1.7fc0bp-26f, 0x1.d39522p-31f, 0.0f, 0x1.20c764p+5f, 0.0f},
    {0x1.4044d8p-13f, -0x1.5c426cp-17f, 0x1.091516p-21f, -0x1.59ef76p-26f, 0x1.9da95cp-31f, 0.0f, 0x1.264788p+5f, 0.0f},
    {0x1.31c78p-13f, -0x1.466724p-17f, 0x1.e7c702p-22f, -0x1.387134p-26f, 0x1.6ec7f8p-31f, 0.0f, 0x1.2bc7acp+5f, 0.0f},
    {0x1.2430eap-13f, -0x1.324756p-17f, 0x1.c1756ep-22f, -0x1.1ab68cp-26f, 0x1.45eb62p-31f, 0.0f, 0x1.3147dp+5f, 0.0f},
    {0x1.176ef4p-13f, -0x1.1fb8dp-17f, 0x1.9ec15ep-22f, -0x1.0044cap-26f, 0x1.223798p-31f, 0.0f, 0x1.36c7f4p+5f, 0.0f},
    {0x1.0b713p-13f, -0x1.0e9604p-17f, 0x1.7f4564p-22f, -0x1.d16538p-27f, 0x1.02f304p-31f, 0.0f, 0x1.3c4818p+5f, 0.0f},
    {0x1.0028bcp-13f, -0x1.fd7af4p-18f, 0x1.62a8b2p-22f, -0x1.a74ab6p-27f, 0x1.cf0158p-32f, 0.0f, 0x1.41c83cp+5f, 0.0f},
    {0x1.eb1018p-14f, -0x1.e022a6p-18f, 0x1.489d5p-22f, -0x1.819da8p-27f, 0x1.9eb6c8p-32f, 0.0f, 0x1.47486p+5f, 0.0f},
    {0x1.d7059ap-14f, -0x1.c4ed9ep-18f, 0x1.30deaep-22f, -0x1.5fd6c8p-27f, 0x1.7423d4p-32f, 0.0f, 0x1.4cc884p+5f, 0.0f},
    {0x1.c41b88p-14f, -0x1.ababe8p-18f, 0x1.1b3054p-22f, -0x1.4180dp-27f, 0x1.4e875p-32f, 0.0f, 0x1.5248a8p+5f, 0.0f},
    {0x1.b23d5cp-14f, -0x1.943278p-18f, 0x1.075cdcp-22f, -0x1.2635d2p-27f, 0x1.2d3c16p-32f, 0.0f, 0x1.57c8ccp+5f, 0.0f},
    {0x1.a15854p-14f, -0x1.7e5a8cp-18f, 0x1.ea69fep-23f, -0x1.0d9d0cp-27f, 0x1.0fb498p-32f, 0.0f, 0x1.5d48fp+5f, 0.0f},
};

__device__ __forceinline__ void near32t(float r2, float sig2, float nu2, const float4* __restrict__ tab,
                                        float& qp, float& fp, float& c) {
  const float isg = __builtin_amdgcn_rsqf(sig2);
  const float isg2 = isg * isg;
  const float isg3 = isg2 * isg;
  const float sb = r2 * isg2;
  const float e = __builtin_amdgcn_exp2f(kMHalfLog2e * sb);
  // fminf keeps the index in the table for any sb (NaN included)
  const int k = static_cast<int>(fminf(sb * kCoreScale, static_cast<float>(kCorePieces - 1)));
  const float4 lo = tab[2 * k], hi = tab[2 * k + 1];  // (c0, c1, c2, c3), (c4, 0, s_c, 0)
  const float t = sb - hi.z;
  float g = fmaf(hi.x, t, lo.w);
  g = fmaf(g, t, lo.z);
  g = fmaf(g, t, lo.y);
  g = fmaf(g, t, lo.x);
  const float ke = kFourPiK * isg3;
  qp = ke * fmaf(sb, g, e * 0x1.555556p-2f);
  fp = -3.0f * (ke * g * isg2);  // -3 f', as accumulate32_pair_n3 takes it
  c = nu2 * kKf * e * isg3 * isg2;  // 2 nu Z / sigma^2
}

// One pair's velocity and stretching into the partial sums (v = a_s x d),
// the stretching factored (A and B).  Used where the source is beyond the
// Gaussian core of every lane's target: there q'|a_s| is small against the
// core pairs'.
template <typename S>
__device__ __forceinline__ void accumulate32(float dx, float dy, float dz, float q, float f, float tax,
                                             float tay, float taz, const S& s, P32& p) {
  const float vx = s.ay * dz - s.az * dy;
  const float vy = s.az * dx - s.ax * dz;
  const float vz = s.ax * dy - s.ay * dx;
  p.ux = fmaf(q, vx, p.ux);
  p.uy = fmaf(q, vy, p.uy);
  p.uz = fmaf(q, vz, p.uz);
  p.ax = fmaf(q, s.ax, p.ax);
  p.ay = fmaf(q, s.ay, p.ay);
  p.az = fmaf(q, s.az, p.az);
  const float g = f * fmaf(taz, vz, fmaf(tay, vy, tax * vx));
  p.bx = fmaf(g, dx, p.bx);
  p.by = fmaf(g, dy, p.by);
  p.bz = fmaf(g, dz, p.bz);
}

// The same sums with the stretching formed per pair: q' (a_t x a_s) goes
// into B as -(q'/3)(a_t x a_s), so that fold() still gives a_t x A - 3 B.
// Used for a source that is inside the Gaussian core of some lane's target:
// there q' is large, A would gather large, nearly parallel strengths (thin
// rings) whose cross product with a_t cancels, and forming the small
// a_t x a_s pair by pair keeps that cancellation out of the sums.
// accumulate32_pair with B kept as fold() would return it: the caller passes
// f3 = -3 f', so q' (a_t x a_s) enters B unscaled and fold_n3() adds the sums
// as they are (B_n3 = sum q' (a_t x a_s) - 3 f' (a_t.v) d = -3 B; one
// multiply less per pair)
template <typename S>
__device__ __forceinline__ void accumulate32_pair_n3(float dx, float dy, float dz, float q, float f3, float tax,
                                                     float tay, float taz, const S& s, P32& p) {
  const float vx = s.ay * dz - s.az * dy;
  const float vy = s.az * dx - s.ax * dz;
  const float vz = s.ax * dy - s.ay * dx;
  p.ux = fmaf(q, vx, p.ux);
  p.uy = fmaf(q, vy, p.uy);
  p.uz = fmaf(q, vz, p.uz);
  const float g = f3 * fmaf(taz, vz, fmaf(tay, vy, tax * vx));
  p.bx = fmaf(q, tay * s.az - taz * s.ay, fmaf(g, dx, p.bx));
  p.by = fmaf(q, taz * s.ax - tax * s.az, fmaf(g, dy, p.by));
  p.bz = fmaf(q, tax * s.ay - tay * s.ax, fmaf(g, dz, p.bz));
}

template <typename S>
__device__ __forceinline__ void accumulate32_pair(float dx, float dy, float dz, float q, float f, float tax,
                                                  float tay, float taz, const S& s, P32& p) {
  const float vx = s.ay * dz - s.az * dy;
  const float vy = s.az * dx - s.ax * dz;
  const float vz = s.ax * dy - s.ay * dx;
  p.ux = fmaf(q, vx, p.ux);
  p.uy = fmaf(q, vy, p.uy);
  p.uz = fmaf(q, vz, p.uz);
  const float g = f * fmaf(taz, vz, fmaf(tay, vy, tax * vx));
  const float m = -0x1.555556p-2f * q;  // -q'/3
  p.bx = fmaf(m, tay * s.az - taz * s.ay, fmaf(g, dx, p.bx));
  p.by = fmaf(m, taz * s.ax - tax * s.az, fmaf(g, dy, p.by));
  p.bz = fmaf(m, tax * s.ay - tay * s.ax, fmaf(g, dz, p.bz));
}

template <typename S>
__device__ __forceinline__ void pair32(const T32& t, const S& s, P32& p, PSE32& e) {
  const float dx = static_cast<float>(t.x - s.x);
  const float dy = static_cast<float>(t.y - s.y);
  const float dz = static_cast<float>(t.z - s.z);
  const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  const float sig2 = s.hs2 + t.hs2;
  const float ir = __builtin_amdgcn_rsqf(r2);
  const float ir2 = ir * ir;
  float q = ir * ir2;
  float f = q * ir2;
  const bool core = r2 <= kFar32 * sig2;
  if (core) {
    float c;
    near32(r2, sig2, t.nu2, q, f, c);
    e.x = fmaf(c, t.vol * s.ax - s.vol * t.ax, e.x);
    e.y = fmaf(c, t.vol * s.ay - s.vol * t.ay, e.y);
    e.z = fmaf(c, t.vol * s.az - s.vol * t.az, e.z);
  }
  // wave-uniform choice: per pair if the source is in any lane's core
  if (__ballot(core))
    accumulate32_pair(dx, dy, dz, q, f, t.ax, t.ay, t.az, s, p);
  else
    accumulate32(dx, dy, dz, q, f, t.ax, t.ay, t.az, s, p);
}

// FP32C pair with the high-order algebraic kernel (highorder.h:20-67), in
// the scaling of the Gaussian FP32 pair (q' = 4 pi q, f' = -(4 pi/3) F):
// with b = rho^2 + sigma^2, q' = (rho^2 + 5/2 sigma^2) b^-5/2 and
// f' = (rho^2 + 7/2 sigma^2) b^-7/2, both exact for every pair.  The PSE
// term 2 nu eta / sigma^2 = 2 nu (105/8pi) sigma^4 b^-9/2 is summed for the
// pairs within 44 sigma^2 (beyond, each term is below 4e-8 of the nearest
// one and all of them together about 1e-6 of the sum), as the Gaussian pair
// does; so is the per-pair stretching choice.
constexpr float kHoEta32 = 0x1.0b615cp+2f;  // 105/(8 pi)

template <typename S>
__device__ __forceinline__ void pair32_ho(const T32& t, const S& s, P32& p, PSE32& e) {
  const float dx = static_cast<float>(t.x - s.x);
  const float dy = static_cast<float>(t.y - s.y);
  const float dz = static_cast<float>(t.z - s.z);
  const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  const float sig2 = s.hs2 + t.hs2;
  const float ib = __builtin_amdgcn_rsqf(r2 + sig2);
  const float ib2 = ib * ib;
  const float ib5 = ib2 * ib2 * ib;
  const float ib7 = ib5 * ib2;
  const float q = fmaf(2.5f, sig2, r2) * ib5;
  const float f = fmaf(3.5f, sig2, r2) * ib7;
  const bool core = r2 <= kFar32 * sig2;
  if (core) {
    const float c = t.nu2 * kHoEta32 * (sig2 * sig2) * (ib7 * ib2);
    e.x = fmaf(c, t.vol * s.ax - s.vol * t.ax, e.x);
    e.y = fmaf(c, t.vol * s.ay - s.vol * t.ay, e.y);
    e.z = fmaf(c, t.vol * s.az - s.vol * t.az, e.z);
  }
  if (__ballot(core))
    accumulate32_pair(dx, dy, dz, q, f, t.ax, t.ay, t.az, s, p);
  else
    accumulate32(dx, dy, dz, q, f, t.ax, t.ay, t.az, s, p);
}

// Near-field source record of the tree path: position relative to the
// target leaf's centre, rounded to FP32 once per particle (for sources beyond
// 44 sigma^2 the rounding, about 6e-8 of the distance to the centre, is a
// relative displacement error below 1e-6, and the pair needs no FP64).
struct alignas(16) SrcRel {
  float rx, ry, rz, ax;  // what the far-only loop reads first: 36 bytes
  float ay, az, mx, my;  // (m = a x r, for the 26-instruction sums)
  float mz, hs2, vol, pad;
};

// Direct-sum source record (64 B): what the separated loop reads first
// (position relative to the block's targets' centre, strength and m = a x r:
// 36 bytes in three broadcast reads), the radius, then the FP64 position for
// the other loops.  The volume, read by those loops only, sits in an array
// of its own.
struct alignas(16) SrcT2 {
  float rx, ry, rz, ax;
  float ay, az, mx, my;
  float mz, hs2;
  double x, y, z;
};
static_assert(sizeof(SrcT2) == 64, "one 64-byte record per source");

// Knuth's two-sum: (hi, lo) += p with the rounding error of hi + p kept.
__device__ __forceinline__ void two_sum(float& hi, float& lo, float p) {
  const float s = hi + p;
  const float bb = s - hi;
  const float err = (hi - (s - bb)) + (p - bb);
  hi = s;
  lo += err;
}

// A chunk's partial sums into the compensated totals (u, then w).
// fold() of sums formed by accumulate32_pair_n3 alone (A = 0, B already -3 B)
__device__ __forceinline__ void fold_n3(const P32& p, float* hi, float* lo) {
  two_sum(hi[0], lo[0], p.ux);
  two_sum(hi[1], lo[1], p.uy);
  two_sum(hi[2], lo[2], p.uz);
  two_sum(hi[3], lo[3], p.bx);
  two_sum(hi[4], lo[4], p.by);
  two_sum(hi[5], lo[5], p.bz);
}

__device__ __forceinline__ void fold(const P32& p, float tax, float tay, float taz, float* hi, float* lo) {
  two_sum(hi[0], lo[0], p.ux);
  two_sum(hi[1], lo[1], p.uy);
  two_sum(hi[2], lo[2], p.uz);
  two_sum(hi[3], lo[3], fmaf(tay, p.az, fmaf(-taz, p.ay, -3.0f * p.bx)));
  two_sum(hi[4], lo[4], fmaf(taz, p.ax, fmaf(-tax, p.az, -3.0f * p.by)));
  two_sum(hi[5], lo[5], fmaf(tax, p.ay, fmaf(-tay, p.ax, -3.0f * p.bz)));
}

// Sums of a separated source range, arranged so that a pair costs 26 vector
// instructions instead of 29.  Positions are relative to a common centre c
// (t' for the target, r_s for the source); with m_s = a_s x r_s (per source)
// and c_t = t' x a_t (per target),
//   a_t . (a_s x d) = a_s . c_t - a_t . m_s,
//   sum q' (a_s x d) = A x t' - M,   A = sum q' a_s,  M = sum q' m_s,
// so a pair forms no cross product of its own.  The cancellation this adds
// is bounded by the separation tests that admit a range (the distances to c
// are within a few pair distances).
struct P32S {
  float ax = 0, ay = 0, az = 0;  // A = sum q' a_s
  float mx = 0, my = 0, mz = 0;  // M = sum q' m_s
  float bx = 0, by = 0, bz = 0;  // B = sum f' (a_t . v) d, v = a_s x d
};

// one pair of a separated chunk: (tx, ty, tz) = t', (cx, cy, cz) = t' x a_t
template <bool HighOrder, typename S>
__device__ __forceinline__ void pair32s(float tx, float ty, float tz, float th2, float tax, float tay, float taz,
                                        float cx, float cy, float cz, const S& s, P32S& p) {
  const float dx = tx - s.rx;
  const float dy = ty - s.ry;
  const float dz = tz - s.rz;
  const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  float q, f;
  if (HighOrder) {
    const float sig2 = s.hs2 + th2;
    const float ib = __builtin_amdgcn_rsqf(r2 + sig2);
    const float ib2 = ib * ib;
    const float ib5 = ib2 * ib2 * ib;
    q = fmaf(2.5f, sig2, r2) * ib5;
    f = fmaf(3.5f, sig2, r2) * (ib5 * ib2);
  } else {
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float ir2 = ir * ir;
    q = ir * ir2;
    f = q * ir2;
  }
  p.ax = fmaf(q, s.ax, p.ax);
  p.ay = fmaf(q, s.ay, p.ay);
  p.az = fmaf(q, s.az, p.az);
  p.mx = fmaf(q, s.mx, p.mx);
  p.my = fmaf(q, s.my, p.my);
  p.mz = fmaf(q, s.mz, p.mz);
  const float h = fmaf(s.az, cz, fmaf(s.ay, cy, fmaf(s.ax, cx, -fmaf(taz, s.mz, fmaf(tay, s.my, tax * s.mx)))));
  const float g = f * h;
  p.bx = fmaf(g, dx, p.bx);
  p.by = fmaf(g, dy, p.by);
  p.bz = fmaf(g, dz, p.bz);
}

// a separated chunk's sums into the compensated totals: u = A x t' - M,
// w = a_t x A - 3 B
__device__ __forceinline__ void fold_s(const P32S& p, float tx, float ty, float tz, float tax, float tay, float taz,
                                       float* hi, float* lo) {
  two_sum(hi[0], lo[0], fmaf(p.ay, tz, fmaf(-p.az, ty, -p.mx)));
  two_sum(hi[1], lo[1], fmaf(p.az, tx, fmaf(-p.ax, tz, -p.my)));
  two_sum(hi[2], lo[2], fmaf(p.ax, ty, fmaf(-p.ay, tx, -p.mz)));
  two_sum(hi[3], lo[3], fmaf(tay, p.az, fmaf(-taz, p.ay, -3.0f * p.bx)));
  two_sum(hi[4], lo[4], fmaf(taz, p.ax, fmaf(-tax, p.az, -3.0f * p.by)));
  two_sum(hi[5], lo[5], fmaf(tax, p.ay, fmaf(-tay, p.ax, -3.0f * p.bz)));
}

// Read-only data of a launch seen through the constant address space: loads
// at wave-uniform addresses then become scalar loads.
template <typename T>
using ConstPtr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ ConstPtr<T> as_const(const T* p) {
  return (ConstPtr<T>)(p);
}
// field by field (a struct copy out of that address space has no constructor)
__device__ __forceinline__ Chunk32 load(ConstPtr<Chunk32> p, int i) {
  Chunk32 h;
  h.cx = p[i].cx; h.cy = p[i].cy; h.cz = p[i].cz;
  h.r = p[i].r; h.h2 = p[i].h2;
  h.first = p[i].first; h.count = p[i].count;
  return h;
}
__device__ __forceinline__ Src32 load(ConstPtr<Src32> p, int i) {
  Src32 s;
  s.rx = p[i].rx; s.ry = p[i].ry; s.rz = p[i].rz;
  s.ax = p[i].ax; s.ay = p[i].ay; s.az = p[i].az;
  s.mx = p[i].mx; s.my = p[i].my; s.mz = p[i].mz;
  s.hs2 = p[i].hs2; s.vol = p[i].vol; s.pad = 0.0f;
  return s;
}

// a source of a range that is not separated, as the core-capable pair reads
// it: FP64 position, the rest in FP32
struct SrcCore {
  double x, y, z;
  float hs2, ax, ay, az, vol;
};

// Bounding boxes of each wave's 64 targets and of each 64-source chunk of a
// staged tile decide, per (wave, chunk), whether every pair is far enough
// apart for the FP32 relative-position loop.  Positions are relative to the
// block's centre c_T (targets within r_T of it), rounded once.  The gap g
// between the wave's and the chunk's bounding spheres must exceed sqrt(44)
// sigma (no pair needs the Gaussian kernel), and
//   r_T + |c_S - c_T| + r_S <= kSepFP32 g,
// which bounds the rounding of the two relative positions, at most 2^-24 of
// their distances to c_T, by 2^-24 kSepFP32 of the displacement (2.4e-7,
// the order of the FP32 pair arithmetic itself).  On the 1M helical wake
// 97% of (wave, chunk) pairs pass, against 79% of (block, tile) pairs
// under the same bound.
constexpr float kSepFP32 = 4.0f;

// HighOrder: the algebraic kernel (pair32_ho, and pair32_rel_ho on the
// separated tiles, which need the same gap as the Gaussian's: it keeps the
// relative positions' rounding small and the PSE terms beyond it are
// dropped in both kernels).
// Comp = false sums every pair of the item into one plain FP32 accumulator
// (no chunks, no two-sum): the uncompensated sum, kept to measure what the
// compensation buys (pw_set_compensation).
template <bool HighOrder, bool Comp = true>
__device__ __forceinline__ void item_fp32c(const PairArgs& pa, const WorkItem& it, SrcT2* tile) {
  __shared__ double s_tb[4][6];
  __shared__ float s_b[4][8];
  __shared__ float s_vol[kTB];
  const Rec64* __restrict__ state = pa.state;
  const float nu = static_cast<float>(pa.nu);
  const int t = it.t0 + static_cast<int>(threadIdx.x);
  const bool active = t < it.t1;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const Rec64 m = state[active ? t : it.t0];
  T32 me;
  me.x = m.x; me.y = m.y; me.z = m.z;
  me.hs2 = static_cast<float>(0.5 * m.s2);
  me.ax = static_cast<float>(m.ax);
  me.ay = static_cast<float>(m.ay);
  me.az = static_cast<float>(m.az);
  me.vol = static_cast<float>(m.vol);
  me.nu2 = 2.0f * nu;
  // the targets' bounding box (inactive lanes repeat the first target)
  {
    const double v[6] = {wave_mind(m.x), wave_mind(m.y), wave_mind(m.z),
                         wave_maxd(m.x), wave_maxd(m.y), wave_maxd(m.z)};
    __syncthreads();
    if (lane == 0)
      for (int k = 0; k < 6; ++k) s_tb[wave][k] = v[k];
  }
  __syncthreads();
  double cT[3];
  float rT2 = 0.0f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double lo = fmin(fmin(s_tb[0][k], s_tb[1][k]), fmin(s_tb[2][k], s_tb[3][k]));
    const double hi = fmax(fmax(s_tb[0][3 + k], s_tb[1][3 + k]), fmax(s_tb[2][3 + k], s_tb[3][3 + k]));
    cT[k] = uniform(0.5 * (lo + hi));
    const float h = static_cast<float>(0.5 * (hi - lo));
    rT2 += h * h;
  }
  const float rT = uniform(sqrtf(rT2) * 1.00001f);
  const float rtx = static_cast<float>(m.x - cT[0]), rty = static_cast<float>(m.y - cT[1]),
              rtz = static_cast<float>(m.z - cT[2]);
  const float h2T = uniform(wave_max(me.hs2));
  // this wave's own targets: sphere relative to c_T
  float cW[3], rW2 = 0.0f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double lo = s_tb[wave][k], hi = s_tb[wave][3 + k];
    cW[k] = uniform(static_cast<float>(0.5 * (lo + hi) - cT[k]));
    const float h = static_cast<float>(0.5 * (hi - lo));
    rW2 += h * h;
  }
  const float rW = uniform(sqrtf(rW2) * 1.00001f);
  // c_t = t' x a_t for the separated ranges
  const float ctx = rty * me.az - rtz * me.ay, cty = rtz * me.ax - rtx * me.az, ctz = rtx * me.ay - rty * me.ax;
  float hi[kAcc], lo[kAcc];
#pragma unroll
  for (int k = 0; k < kAcc; ++k) hi[k] = lo[k] = 0.0f;
  PSE32 pse;
  P32 whole;     // Comp == false: the item's plain FP32 sums (one frame: the
  P32S whole_s;  // separated ones fold once, at the end)
  for (int base = it.s0; base < it.s1; base += kTB) {
    const int cnt = min(kTB, it.s1 - base);
    __syncthreads();
    float bl[3] = {INFINITY, INFINITY, INFINITY}, bh[3] = {-INFINITY, -INFINITY, -INFINITY}, h2 = 0.0f;
    if (static_cast<int>(threadIdx.x) < cnt) {
      const Rec64 r = state[base + threadIdx.x];
      SrcT2 f;
      f.rx = static_cast<float>(r.x - cT[0]);
      f.ry = static_cast<float>(r.y - cT[1]);
      f.rz = static_cast<float>(r.z - cT[2]);
      f.hs2 = static_cast<float>(0.5 * r.s2);
      f.ax = static_cast<float>(r.ax);
      f.ay = static_cast<float>(r.ay);
      f.az = static_cast<float>(r.az);
      f.mx = f.ay * f.rz - f.az * f.ry;
      f.my = f.az * f.rx - f.ax * f.rz;
      f.mz = f.ax * f.ry - f.ay * f.rx;
      f.x = r.x; f.y = r.y; f.z = r.z;
      tile[threadIdx.x] = f;
      s_vol[threadIdx.x] = static_cast<float>(r.vol);
      bl[0] = bh[0] = f.rx;
      bl[1] = bh[1] = f.ry;
      bl[2] = bh[2] = f.rz;
      h2 = f.hs2;
    }
    {
      const float v[7] = {wave_min(bl[0]), wave_min(bl[1]), wave_min(bl[2]), wave_max(bh[0]),
                          wave_max(bh[1]), wave_max(bh[2]), wave_max(h2)};
      if (lane == 0)
        for (int k = 0; k < 7; ++k) s_b[wave][k] = v[k];
    }
    __syncthreads();
    if (!active) continue;
    const bool diag = base == it.t0;
    const int self = static_cast<int>(threadIdx.x);
    static_assert(kChunk == 64 && kTB == 4 * kChunk, "a chunk is the sources one wave staged");
    for (int c0 = 0; c0 < cnt; c0 += kChunk) {
      const int c1 = min(cnt, c0 + kChunk);
      const int cw = c0 / kChunk;
      // the chunk's sphere (relative to c_T) against this wave's
      float cs2 = 0.0f, rs2 = 0.0f, d2 = 0.0f;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float l = s_b[cw][k], h = s_b[cw][3 + k];
        const float c = 0.5f * (l + h), e = 0.5f * (h - l);
        cs2 += c * c;
        rs2 += e * e;
        d2 += (c - cW[k]) * (c - cW[k]);
      }
      const float h2S = uniform(s_b[cw][6]);
      const float rS = uniform(sqrtf(rs2) * 1.00001f);
      const float gap = uniform(sqrtf(d2) * 0.99999f - rW - rS);
      const bool sep = gap > 0.0f && gap * gap > kFar32 * (h2T + h2S) * 1.0001f &&
                       rT + uniform(sqrtf(cs2) * 1.00001f) + rS <= kSepFP32 * gap;
      if (sep) {
        P32S chunk_s;
        P32S& q = Comp ? chunk_s : whole_s;
#pragma unroll 2
        for (int l = c0; l < c1; ++l)
          pair32s<HighOrder>(rtx, rty, rtz, me.hs2, me.ax, me.ay, me.az, ctx, cty, ctz, tile[l], q);
        if (Comp) fold_s(q, rtx, rty, rtz, me.ax, me.ay, me.az, hi, lo);
        continue;
      }
      P32 chunk;
      P32& p = Comp ? chunk : whole;
      const bool own = diag && cw == wave;  // the chunk holding this wave's own targets
#pragma unroll 1
      for (int l = c0; l < c1; ++l) {
        if (own && l == self) continue;  // the self pair is never evaluated
        const SrcT2& f = tile[l];
        SrcCore sc;
        sc.x = f.x; sc.y = f.y; sc.z = f.z;
        sc.hs2 = f.hs2; sc.ax = f.ax; sc.ay = f.ay; sc.az = f.az; sc.vol = s_vol[l];
        if (HighOrder) pair32_ho(me, sc, p, pse);
        else pair32(me, sc, p, pse);
      }
      if (Comp) fold(p, me.ax, me.ay, me.az, hi, lo);
    }
  }
  if (!Comp) {
    fold(whole, me.ax, me.ay, me.az, hi, lo);
    fold_s(whole_s, rtx, rty, rtz, me.ax, me.ay, me.az, hi, lo);
  }
  if (active) {
    double acc[kAcc];
#pragma unroll
    for (int k = 0; k < kAcc; ++k) acc[k] = static_cast<double>(hi[k]) + static_cast<double>(lo[k]);
    Rate64 r;
    r.ux = kQ * acc[0];
    r.uy = kQ * acc[1];
    r.uz = kQ * acc[2];
    r.wx = kQ * acc[3] + static_cast<double>(pse.x);
    r.wy = kQ * acc[4] + static_cast<double>(pse.y);
    r.wz = kQ * acc[5] + static_cast<double>(pse.z);
    out_rates(pa, it)[t] = r;
  }
}

// ---------------------------------------------------------------------
// FP32C direct sum, scalar-fed (PairConfig::scalar_feed).  One wave owns 64
// targets of a work item and streams the item's sources chunk by chunk
// straight from memory with scalar loads: a source record is the same for
// every lane, so it sits in scalar registers and enters the vector
// instructions as an operand, with no LDS staging, no per-lane copies and
// no barriers.  k_src_chunks rebuilds the chunks' spheres and the relative
// source records from the state before each launch.
//
// A chunk whose bounding sphere is separated from the wave's targets' (gap
// beyond sqrt(44) sigma, and at least the chunk's radius) runs on positions
// relative to the chunk's centre c: t' = x_t - c (FP64, rounded once per
// chunk) against r_s = x_s - c (rounded once per launch).  The displacement
// t' - r_s then carries a rounding error of at most 2^-24 (2|d| + 2 r_S),
// i.e. 4 * 2^-24 of |d|, the bound of the LDS-fed kernel's separated loop;
// its sums are the 26-instruction ones (pair32s), c_t formed per chunk.  The
// other chunks take the LDS-fed kernel's pair (FP64 displacement, Gaussian
// core, stretching formed per pair for sources inside some lane's core).
// ---------------------------------------------------------------------

// Chunk spheres and relative source records from the state (one wave per
// chunk; first and count were set when the context was built).
__global__ __launch_bounds__(64) void k_src_chunks(const Rec64* __restrict__ state, Chunk32* __restrict__ chunks,
                                                   Src32* __restrict__ src) {
  Chunk32& h = chunks[blockIdx.x];
  const int first = h.first, count = h.count;
  const int lane = static_cast<int>(threadIdx.x);
  const bool on = lane < count;
  const Rec64 r = state[first + (on ? lane : 0)];
  const double lo[3] = {wave_mind(r.x), wave_mind(r.y), wave_mind(r.z)};
  const double hi[3] = {wave_maxd(r.x), wave_maxd(r.y), wave_maxd(r.z)};
  double c[3];
  float e2 = 0.0f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    c[k] = 0.5 * (lo[k] + hi[k]);
    const float e = static_cast<float>(0.5 * (hi[k] - lo[k]));
    e2 += e * e;
  }
  const float hs2 = static_cast<float>(0.5 * r.s2);
  const float h2 = wave_max(hs2);
  if (lane == 0) {
    h.cx = c[0];
    h.cy = c[1];
    h.cz = c[2];
    h.r = sqrtf(e2) * 1.00001f + 1e-30f;
    h.h2 = h2;
  }
  if (on) {
    Src32 o;
    o.rx = static_cast<float>(r.x - c[0]);
    o.ry = static_cast<float>(r.y - c[1]);
    o.rz = static_cast<float>(r.z - c[2]);
    o.ax = static_cast<float>(r.ax);
    o.ay = static_cast<float>(r.ay);
    o.az = static_cast<float>(r.az);
    o.mx = o.ay * o.rz - o.az * o.ry;
    o.my = o.az * o.rx - o.ax * o.rz;
    o.mz = o.ax * o.ry - o.ay * o.rx;
    o.hs2 = hs2;
    o.vol = static_cast<float>(r.vol);
    o.pad = 0.0f;
    src[first + lane] = o;
  }
}

// Targets [tb, tb + 64 TPL) of item it (a 64-lane wave of its own; with
// TPL = 2 each lane owns targets tb + lane and tb + 64 + lane, and every
// source record read from scalar registers feeds two pairs).
template <bool HighOrder, int TPL>
__device__ __forceinline__ void wave_fp32s(const WorkItem& it, int tb, double nu, const Rec64* state_g,
                                           const Chunk32* chunks_g, int nchunks, const Src32* src_g,
                                           Rate64* __restrict__ out) {
  const ConstPtr<Rec64> state = as_const(state_g);
  const ConstPtr<Chunk32> chunks = as_const(chunks_g);
  const ConstPtr<Src32> src = as_const(src_g);
  const int lane = static_cast<int>(threadIdx.x) & 63;
  int t[TPL];
  bool active[TPL];
  T32 me[TPL];
  double lo_x = 0, lo_y = 0, lo_z = 0, hi_x = 0, hi_y = 0, hi_z = 0;
  float hs2max = 0.0f;
#pragma unroll
  for (int j = 0; j < TPL; ++j) {
    t[j] = tb + 64 * j + lane;
    active[j] = t[j] < it.t1;
    // inactive lanes repeat the first target (inside the sphere already)
    const Rec64 m = state_g[active[j] ? t[j] : tb];
    me[j].x = m.x; me[j].y = m.y; me[j].z = m.z;
    me[j].hs2 = static_cast<float>(0.5 * m.s2);
    me[j].ax = static_cast<float>(m.ax);
    me[j].ay = static_cast<float>(m.ay);
    me[j].az = static_cast<float>(m.az);
    me[j].vol = static_cast<float>(m.vol);
    me[j].nu2 = 2.0f * static_cast<float>(nu);
    lo_x = j ? fmin(lo_x, m.x) : m.x; hi_x = j ? fmax(hi_x, m.x) : m.x;
    lo_y = j ? fmin(lo_y, m.y) : m.y; hi_y = j ? fmax(hi_y, m.y) : m.y;
    lo_z = j ? fmin(lo_z, m.z) : m.z; hi_z = j ? fmax(hi_z, m.z) : m.z;
    hs2max = fmaxf(hs2max, me[j].hs2);
  }
  // the wave's targets' sphere
  double cW[3];
  float rW2 = 0.0f;
  {
    const double lo[3] = {wave_mind(lo_x), wave_mind(lo_y), wave_mind(lo_z)};
    const double hi[3] = {wave_maxd(hi_x), wave_maxd(hi_y), wave_maxd(hi_z)};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      cW[k] = uniform(0.5 * (lo[k] + hi[k]));
      const float e = static_cast<float>(0.5 * (hi[k] - lo[k]));
      rW2 += e * e;
    }
  }
  const float rW = uniform(sqrtf(rW2) * 1.00001f);
  const float h2W = uniform(wave_max(hs2max));
  float hi[TPL][kAcc], lo[TPL][kAcc];
#pragma unroll
  for (int j = 0; j < TPL; ++j)
#pragma unroll
    for (int k = 0; k < kAcc; ++k) hi[j][k] = lo[j][k] = 0.0f;
  PSE32 pse[TPL];
  for (int ci = it.chunk0; ci < nchunks; ++ci) {
    const Chunk32 h = load(chunks, ci);
    if (h.first >= it.s1) break;
    const int l0 = max(h.first, it.s0), l1 = min(h.first + h.count, it.s1);
    const double ddx = h.cx - cW[0], ddy = h.cy - cW[1], ddz = h.cz - cW[2];
    const float dist = static_cast<float>(sqrt(fma(ddz, ddz, fma(ddy, ddy, ddx * ddx))));
    const float gap = dist * 0.99999f - rW - h.r;
    const bool sep = gap > 0.0f && gap * gap > kFar32 * (h2W + h.h2) * 1.0001f && h.r <= gap;
    if (sep) {
      float tx[TPL], ty[TPL], tz[TPL], cx[TPL], cy[TPL], cz[TPL];
#pragma unroll
      for (int j = 0; j < TPL; ++j) {
        tx[j] = static_cast<float>(me[j].x - h.cx);
        ty[j] = static_cast<float>(me[j].y - h.cy);
        tz[j] = static_cast<float>(me[j].z - h.cz);
        cx[j] = ty[j] * me[j].az - tz[j] * me[j].ay;
        cy[j] = tz[j] * me[j].ax - tx[j] * me[j].az;
        cz[j] = tx[j] * me[j].ay - ty[j] * me[j].ax;
      }
      P32S p[TPL];
      constexpr int kSepUnroll = 4 / TPL;
#pragma unroll kSepUnroll
      for (int l = l0; l < l1; ++l) {
        const Src32 sr = load(src, l);
#pragma unroll
        for (int j = 0; j < TPL; ++j)
          pair32s<HighOrder>(tx[j], ty[j], tz[j], me[j].hs2, me[j].ax, me[j].ay, me[j].az, cx[j], cy[j], cz[j], sr,
                             p[j]);
      }
#pragma unroll
      for (int j = 0; j < TPL; ++j) fold_s(p[j], tx[j], ty[j], tz[j], me[j].ax, me[j].ay, me[j].az, hi[j], lo[j]);
    } else {
      P32 p[TPL];
#pragma unroll 1
      for (int l = l0; l < l1; ++l) {
        const Src32 sr = load(src, l);
        SrcCore sc;
        sc.x = state[l].x; sc.y = state[l].y; sc.z = state[l].z;
        sc.hs2 = sr.hs2; sc.ax = sr.ax; sc.ay = sr.ay; sc.az = sr.az; sc.vol = sr.vol;
#pragma unroll
        for (int j = 0; j < TPL; ++j) {
          if (l != t[j]) {  // the self pair is never evaluated
            if (HighOrder) pair32_ho(me[j], sc, p[j], pse[j]);
            else pair32(me[j], sc, p[j], pse[j]);
          }
        }
      }
#pragma unroll
      for (int j = 0; j < TPL; ++j) fold(p[j], me[j].ax, me[j].ay, me[j].az, hi[j], lo[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < TPL; ++j) {
    if (!active[j]) continue;
    double acc[kAcc];
#pragma unroll
    for (int k = 0; k < kAcc; ++k) acc[k] = static_cast<double>(hi[j][k]) + static_cast<double>(lo[j][k]);
    Rate64 r;
    r.ux = kQ * acc[0];
    r.uy = kQ * acc[1];
    r.uz = kQ * acc[2];
    r.wx = kQ * acc[3] + static_cast<double>(pse[j].x);
    r.wy = kQ * acc[4] + static_cast<double>(pse[j].y);
    r.wz = kQ * acc[5] + static_cast<double>(pse[j].z);
    out[t[j]] = r;
  }
}

// Entry: one-wave blocks; wave index w = (4 / TPL) item + part of the item
// (grid form: w = blockIdx.x; persistent form: pulled from the queue).
template <bool HighOrder, bool Persistent, int TPL = 1>
__device__ __forceinline__ void fp32s_entry(const PairArgs& a) {
  constexpr int kWavesPerItem = 4 / TPL;
  const int waves = kWavesPerItem * a.nitems;
  for (int w = Persistent ? -1 : static_cast<int>(blockIdx.x);;) {
    if (Persistent) {
      int idx = 0;
      if (threadIdx.x == 0) idx = atomicAdd(a.queue, 1);
      w = __builtin_amdgcn_readfirstlane(idx);  // lane 0's: every lane is active here
    }
    if (w >= waves) break;
    const WorkItem it = a.items[w / kWavesPerItem];
    const int tb = it.t0 + 64 * TPL * (w % kWavesPerItem);
    if (tb < it.t1) {
      Rate64* out = it.part < 0 ? a.rates : a.partial + static_cast<size_t>(it.part) * a.n;
      wave_fp32s<HighOrder, TPL>(it, tb, a.nu, a.state, a.chunks, a.nchunks, a.src, out);
    }
    if (!Persistent) break;
  }
}

// ---------------------------------------------------------------------
// FP64 direct sum, scalar-fed (PairConfig::scalar_feed for PW_FP64): the
// FP32C scalar-fed scheme with FP64 records.  One wave owns 64 targets and
// streams its item's sources chunk by chunk through scalar loads; a chunk
// separated from the wave's targets (gap beyond sqrt(90) sigma, where the
// Gaussian terms are below FP64 rounding, and at least the chunk's radius)
// runs the separated sums of pair64s on positions relative to the chunk's
// centre (t' rounded once per chunk, r_s once per launch, both FP64: the
// displacement's rounding is at most 2^-53 (2|d| + 2 r_S) <= 4 x 2^-53 of
// |d|), the others the LDS kernel's pair64 on FP64 positions.  The 72 bytes
// a separated pair reads would be 4.6 KB of LDS traffic per wave and pair;
// from scalar registers they cost no vector-memory or LDS bandwidth.
// ---------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_src_chunks64(const Rec64* __restrict__ state, Chunk32* __restrict__ chunks,
                                                     Src64* __restrict__ src) {
  Chunk32& h = chunks[blockIdx.x];
  const int first = h.first, count = h.count;
  const int lane = static_cast<int>(threadIdx.x);
  const bool on = lane < count;
  const Rec64 r = state[first + (on ? lane : 0)];
  const double lo[3] = {wave_mind(r.x), wave_mind(r.y), wave_mind(r.z)};
  const double hi[3] = {wave_maxd(r.x), wave_maxd(r.y), wave_maxd(r.z)};
  double c[3], e2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    c[k] = 0.5 * (lo[k] + hi[k]);
    e2 += 0.25 * (hi[k] - lo[k]) * (hi[k] - lo[k]);
  }
  const float h2 = static_cast<float>(wave_maxd(0.5 * r.s2) * 1.000001);
  if (lane == 0) {
    h.cx = c[0];
    h.cy = c[1];
    h.cz = c[2];
    h.r = static_cast<float>(sqrt(e2) * 1.000001) + 1e-30f;
    h.h2 = h2;
  }
  if (on) {
    Src64 o;
    o.rx = r.x - c[0];
    o.ry = r.y - c[1];
    o.rz = r.z - c[2];
    o.ax = r.ax;
    o.ay = r.ay;
    o.az = r.az;
    o.mx = r.ay * o.rz - r.az * o.ry;
    o.my = r.az * o.rx - r.ax * o.rz;
    o.mz = r.ax * o.ry - r.ay * o.rx;
    o.pad = 0.0;
    src[first + lane] = o;
  }
}

__device__ __forceinline__ Rec64 load(ConstPtr<Rec64> p, int i) {
  Rec64 r;
  r.x = p[i].x; r.y = p[i].y; r.z = p[i].z; r.s2 = p[i].s2;
  r.ax = p[i].ax; r.ay = p[i].ay; r.az = p[i].az; r.vol = p[i].vol;
  return r;
}

// Targets [tb, tb + 64) of item it (a 64-lane wave of its own).
__device__ __forceinline__ void wave_fp64s(const WorkItem& it, int tb, double nu, const Rec64* state_g,
                                           const Chunk32* chunks_g, int nchunks, const Src64* src_g,
                                           Rate64* __restrict__ out) {
  const ConstPtr<Rec64> state = as_const(state_g);
  const ConstPtr<Chunk32> chunks = as_const(chunks_g);
  const ConstPtr<Src64> src = as_const(src_g);
  const int lane = static_cast<int>(threadIdx.x) & 63;
  const int t = tb + lane;
  const bool active = t < it.t1;
  const Rec64 m = state_g[active ? t : tb];
  T64 me;
  me.x = m.x; me.y = m.y; me.z = m.z; me.hs2 = 0.5 * m.s2;
  me.ax = m.ax; me.ay = m.ay; me.az = m.az; me.vol = m.vol;
  me.nu2 = 2.0 * nu;
  // the wave's targets' sphere (inactive lanes repeat the first target)
  double cW[3], rW2 = 0.0;
  {
    const double lo[3] = {wave_mind(m.x), wave_mind(m.y), wave_mind(m.z)};
    const double hi[3] = {wave_maxd(m.x), wave_maxd(m.y), wave_maxd(m.z)};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      cW[k] = uniform(0.5 * (lo[k] + hi[k]));
      rW2 += 0.25 * (hi[k] - lo[k]) * (hi[k] - lo[k]);
    }
  }
  const double rW = uniform(sqrt(rW2) * 1.000001);
  const double h2W = uniform(wave_maxd(me.hs2));
  AccF64 a;
  for (int ci = it.chunk0; ci < nchunks; ++ci) {
    const Chunk32 h = load(chunks, ci);
    if (h.first >= it.s1) break;
    const int l0 = max(h.first, it.s0), l1 = min(h.first + h.count, it.s1);
    const double ddx = h.cx - cW[0], ddy = h.cy - cW[1], ddz = h.cz - cW[2];
    const double gap = sqrt(ddx * ddx + ddy * ddy + ddz * ddz) * 0.999999 - rW - h.r;
    const bool sep = gap > 0.0 && gap * gap > kFar64 * (h2W + h.h2) * 1.0001 && h.r <= gap;
    if (sep) {
      const double tx = me.x - h.cx, ty = me.y - h.cy, tz = me.z - h.cz;
      const double cx = ty * me.az - tz * me.ay, cy = tz * me.ax - tx * me.az, cz = tx * me.ay - ty * me.ax;
      P64S p;
#pragma unroll 2
      for (int l = l0; l < l1; ++l)
        pair64s(tx, ty, tz, me.ax, me.ay, me.az, cx, cy, cz, src[l].rx, src[l].ry, src[l].rz, src[l].ax,
                src[l].ay, src[l].az, src[l].mx, src[l].my, src[l].mz, p);
      // u += (A x t' - M)/4pi; the stretching's A and B join the factored sums
      a.ux += kQ * ((p.ay * tz - p.az * ty) - p.mx);
      a.uy += kQ * ((p.az * tx - p.ax * tz) - p.my);
      a.uz += kQ * ((p.ax * ty - p.ay * tx) - p.mz);
      a.ax += kQ * p.ax;
      a.ay += kQ * p.ay;
      a.az += kQ * p.az;
      a.bx += (-3.0 * kQ) * p.bx;
      a.by += (-3.0 * kQ) * p.by;
      a.bz += (-3.0 * kQ) * p.bz;
    } else {
#pragma unroll 1
      for (int l = l0; l < l1; ++l)
        if (l != t) pair64(me, load(state, l), a);  // the self pair is never evaluated
    }
  }
  if (active) {
    Rate64 r;
    r.ux = a.ux; r.uy = a.uy; r.uz = a.uz;
    r.wx = (me.ay * a.az - me.az * a.ay) + a.bx;
    r.wy = (me.az * a.ax - me.ax * a.az) + a.by;
    r.wz = (me.ax * a.ay - me.ay * a.ax) + a.bz;
    out[t] = r;
  }
}

template <bool Persistent>
__device__ __forceinline__ void fp64s_entry(const PairArgs& a) {
  const int waves = 4 * a.nitems;
  for (int w = Persistent ? -1 : static_cast<int>(blockIdx.x);;) {
    if (Persistent) {
      int idx = 0;
      if (threadIdx.x == 0) idx = atomicAdd(a.queue, 1);
      w = __builtin_amdgcn_readfirstlane(idx);  // lane 0's: every lane is active here
    }
    if (w >= waves) break;
    const WorkItem it = a.items[w >> 2];
    const int tb = it.t0 + 64 * (w & 3);
    if (tb < it.t1) {
      Rate64* out = it.part < 0 ? a.rates : a.partial + static_cast<size_t>(it.part) * a.n;
      wave_fp64s(it, tb, a.nu, a.state, a.chunks, a.nchunks, a.src64, out);
    }
    if (!Persistent) break;
  }
}

// ---------------------------------------------------------------------
// Near field of the tree path.  A work item is (target leaf T, a range of
// T's near list); its targets are cut into 64-lane slices, and one slice of
// one item is what a wave pulls from the atomic queue (queue index =
// 4 item + slice; slices past the leaf's end are skipped).  Each wave is a
// block of its own: it stages the listed leaves' sources 64 at a time into
// its own LDS tile and waits on no other wave, so leaves of any size keep
// every resident wave busy.  A leaf whose list is split writes per-item
// partial sums that k_near_reduce adds in item order (deterministic).
//
// Positions are FP32 relative to T's centre for every listed leaf.  The
// rounding is 2^-24 of the distance to the centre, which for near pairs is
// a few leaf radii: the displacement of the closest pairs carries a relative
// error of about 2^-24 (leaf radius / spacing), some 1e-6 on these wakes,
// far below the error of the far field the near field is paired with.
// Listed leaves whose every pair is beyond 44 sigma^2 skip the Gaussian
// core test.
// ---------------------------------------------------------------------
struct NearArgs {
  const Rec64* sorted;
  const NodeRec* nodes;
  const int* list;
  int cap;
  const NearItem* items;
  const int* total;  // items in use (device)
  int* queue;        // slices handed out
  float nu;
  float gap2;  // 44 sigma_max^2: leaves (or runs) further apart take the far-only loop
  Rate64* out;
  Rate64* partial;  // item x 256
  const LeafChunk* chunks;  // leaf x kLeafChunks run spheres
  // every particle of the same radius (uniform sigma): the core's sigma
  // factors, once per launch (u_cut2 = 44 sigma^2 > 0 marks the case):
  // 1/sigma^2, ke = sqrt(2/pi)/sigma^3 (the table is staged times ke),
  // log2(ke/3) (added to the exponent: e' = (ke/3) e^{-s/2}; staged into
  // the table's rows), -3/sigma^2 (f' = g/sigma^2 enters the sums as -3 f')
  // and the PSE weight 2 nu (2 pi)^-3/2 / sigma^5 divided by ke/3
  float u_cut2, u_isg2, u_ke, u_lk3, u_ck3, u_n3isg2;
};

constexpr int kNearSlice = 64;                       // targets per wave
static_assert(kTB == kLeafSize, "partial-sum rows hold one leaf");
constexpr int kNearSlices = kLeafSize / kNearSlice;  // slices per leaf at most
static_assert(kNearSlices == kLeafChunks && kNearSlice == 64, "a slice is one 64-particle run of its leaf");

// pair on relative FP32 positions, Gaussian core included (HighOrder: the
// algebraic kernel, exact for every pair, PSE within 44 sigma^2 as in
// pair32_ho)
// near32t for a pair radius shared by every pair: the sigma factors come
// from the launch and ke is folded into the staged table (g = ke G) and
// into the exponent (e = (ke/3) e^{-s/2}, log2(ke/3) staged as each row's
// hi.y), so q' = s g + e and -3 f' = -3 g / sigma^2.  Only for r2 <= u_cut2:
// then sb <= fl(u_cut2 / sigma^2) and the index below stays under
// kCorePieces, which the launch checks in FP32 (rounding is monotone).
__device__ __forceinline__ void near32u(float r2, const NearArgs& a, const float4* __restrict__ tab, float& qp,
                                        float& fp3, float& c) {
  const float sb = r2 * a.u_isg2;
  const int k = static_cast<int>(sb * kCoreScale);
  const float4 lo = tab[2 * k], hi = tab[2 * k + 1];  // (c0, c1, c2, c3), (c4, log2(ke/3), s_c, 0)
  const float e = __builtin_amdgcn_exp2f(fmaf(kMHalfLog2e, sb, hi.y));
  const float t = sb - hi.z;
  float g = fmaf(hi.x, t, lo.w);
  g = fmaf(g, t, lo.z);
  g = fmaf(g, t, lo.y);
  g = fmaf(g, t, lo.x);
  qp = fmaf(sb, g, e);
  fp3 = a.u_n3isg2 * g;
  c = a.u_ck3 * e;
}

template <bool HighOrder, bool Uniform = false>
__device__ __forceinline__ void pair32_rel_core(float tx, float ty, float tz, const T32& t, const SrcRel& s,
                                                const float4* __restrict__ tab, const NearArgs& a, P32& p,
                                                PSE32F& e) {
  const float dx = tx - s.rx;
  const float dy = ty - s.ry;
  const float dz = tz - s.rz;
  const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  const float sig2 = Uniform ? 0.0f : s.hs2 + t.hs2;
  float q, f;
  float ib7 = 0.0f, ib2 = 0.0f;
  if (HighOrder) {
    const float ib = __builtin_amdgcn_rsqf(r2 + sig2);
    ib2 = ib * ib;
    const float ib5 = ib2 * ib2 * ib;
    ib7 = ib5 * ib2;
    q = fmaf(2.5f, sig2, r2) * ib5;
    f = fmaf(3.5f, sig2, r2) * ib7;
  } else {  // f holds -3 f' (accumulate32_pair_n3)
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float ir2 = ir * ir;
    q = ir * ir2;
    f = -3.0f * (q * ir2);
  }
  const bool core = Uniform ? r2 <= a.u_cut2 : r2 <= kFar32 * sig2;
  if (core) {
    float c;
    if (HighOrder) c = t.nu2 * kHoEta32 * (sig2 * sig2) * (ib7 * ib2);
    else if (Uniform) near32u(r2, a, tab, q, f, c);
    else near32t(r2, sig2, t.nu2, tab, q, f, c);
    e.x = fmaf(c, s.ax, e.x);
    e.y = fmaf(c, s.ay, e.y);
    e.z = fmaf(c, s.az, e.z);
    e.w = fmaf(c, s.vol, e.w);
  }
  // a run within the core range of the slice: stretching per pair (a_t x a_s
  // formed per pair keeps its accuracy where a_s is nearly parallel to a_t)
  if (HighOrder) accumulate32_pair(dx, dy, dz, q, f, t.ax, t.ay, t.az, s, p);
  else accumulate32_pair_n3(dx, dy, dz, q, f, t.ax, t.ay, t.az, s, p);
}

// Halves: the slice holds at most 32 targets; lanes 32-63 take the same
// targets as lanes 0-31 and each half every other source of a tile, and
// the two halves' sums are added at the end (a leaf's last slice is often
// this short: 13% of the 1M helix's slices)
template <bool HighOrder, bool Uniform, bool Halves>
__device__ __forceinline__ void near_slice(const NearArgs& a, const NearItem& it, int slice, SrcRel* tile,
                                           const float4* __restrict__ tab) {
  const int T = it.T;
  const int lane = static_cast<int>(threadIdx.x);
  const int h = Halves ? lane >> 5 : 0;  // this lane's half: sources h, h + 2, ...
  constexpr int kStep = Halves ? 2 : 1;
  const int t0 = a.nodes[T].first, t1 = a.nodes[T].last;
  const double tcx = a.nodes[T].cx, tcy = a.nodes[T].cy, tcz = a.nodes[T].cz;
  const float tr = a.nodes[T].r;
  const LeafChunk ts = a.chunks[kLeafChunks * T + slice];  // this slice's targets
  const int self = kNearSlice * slice + (Halves ? lane & 31 : lane);  // index of this lane's target in T
  const int t = t0 + self;
  const bool active = t < t1;
  const Rec64 m = a.sorted[active ? t : t0];
  T32 me;
  me.x = m.x; me.y = m.y; me.z = m.z;
  me.hs2 = static_cast<float>(0.5 * m.s2);
  me.ax = static_cast<float>(m.ax);
  me.ay = static_cast<float>(m.ay);
  me.az = static_cast<float>(m.az);
  me.vol = static_cast<float>(m.vol);
  me.nu2 = 2.0f * a.nu;
  const float rx = static_cast<float>(m.x - tcx), ry = static_cast<float>(m.y - tcy),
              rz = static_cast<float>(m.z - tcz);
  // c_t = t' x a_t for the far-only leaves (positions relative to T's centre)
  const float ctx = ry * me.az - rz * me.ay, cty = rz * me.ax - rx * me.az, ctz = rx * me.ay - ry * me.ax;
  float hi[kAcc], lo[kAcc];
#pragma unroll
  for (int k = 0; k < kAcc; ++k) hi[k] = lo[k] = 0.0f;
  PSE32F pse;
  const int* __restrict__ L = a.list + static_cast<size_t>(T) * a.cap;
  for (int k = it.k0; k < it.k1; ++k) {
    const int id = L[k];
    const int base = a.nodes[id].first;
    const int cnt = a.nodes[id].last - base;
    // every pair of the two leaves beyond the Gaussian core: no core test
    const double ddx = a.nodes[id].cx - tcx, ddy = a.nodes[id].cy - tcy, ddz = a.nodes[id].cz - tcz;
    const float gap = static_cast<float>(sqrt(ddx * ddx + ddy * ddy + ddz * ddz)) - tr - a.nodes[id].r;
    const bool all_far = gap > 0.0f && gap * gap > a.gap2 * 1.0001f;
    const bool diag = id == T;
    for (int c0 = 0; c0 < cnt; c0 += kNearSlice) {
      const int nc = min(kNearSlice, cnt - c0);
      __syncthreads();  // the previous tile is read
      if (lane < nc) {
        const Rec64 r = a.sorted[base + c0 + lane];
        SrcRel f;
        f.rx = static_cast<float>(r.x - tcx);
        f.ry = static_cast<float>(r.y - tcy);
        f.rz = static_cast<float>(r.z - tcz);
        f.hs2 = static_cast<float>(0.5 * r.s2);
        f.ax = static_cast<float>(r.ax);
        f.ay = static_cast<float>(r.ay);
        f.az = static_cast<float>(r.az);
        f.mx = f.ay * f.rz - f.az * f.ry;
        f.my = f.az * f.rx - f.ax * f.rz;
        f.mz = f.ax * f.ry - f.ay * f.rx;
        f.vol = static_cast<float>(r.vol);
        f.pad = 0.0f;
        tile[lane] = f;
      }
      __syncthreads();
      if (!active) continue;
      // a run of the leaf beyond the core of every target of the slice
      bool sep = all_far;
      if (!sep) {
        const LeafChunk& cs = a.chunks[kLeafChunks * id + (c0 >> 6)];
        const double ex = cs.cx - ts.cx, ey = cs.cy - ts.cy, ez = cs.cz - ts.cz;
        const float g = static_cast<float>(sqrt(ex * ex + ey * ey + ez * ez)) - ts.r - cs.r;
        sep = g > 0.0f && g * g > a.gap2 * 1.0001f;
      }
      if (sep) {
        P32S q;
#pragma unroll 2
        for (int l = h; l < nc; l += kStep)
          pair32s<HighOrder>(rx, ry, rz, me.hs2, me.ax, me.ay, me.az, ctx, cty, ctz, tile[l], q);
        fold_s(q, rx, ry, rz, me.ax, me.ay, me.az, hi, lo);
        continue;
      }
      P32 p;
      if (diag) {
        const int own = self - c0;  // this target's slot in the tile, if there
        for (int l = h; l < nc; l += kStep)
          if (l != own) pair32_rel_core<HighOrder, Uniform>(rx, ry, rz, me, tile[l], tab, a, p, pse);
      } else {
#pragma unroll 1
        for (int l = h; l < nc; l += kStep) pair32_rel_core<HighOrder, Uniform>(rx, ry, rz, me, tile[l], tab, a, p, pse);
      }
      if (HighOrder) fold(p, me.ax, me.ay, me.az, hi, lo);
      else fold_n3(p, hi, lo);
    }
  }
  if (Halves) {  // lane l (< 32) takes lane l + 32's sums; every lane runs the shuffles
#pragma unroll
    for (int k = 0; k < kAcc; ++k) {
      const float oh = __shfl_down(hi[k], 32), ol = __shfl_down(lo[k], 32);
      two_sum(hi[k], lo[k], oh);
      lo[k] += ol;
    }
    pse.x += __shfl_down(pse.x, 32);
    pse.y += __shfl_down(pse.y, 32);
    pse.z += __shfl_down(pse.z, 32);
    pse.w += __shfl_down(pse.w, 32);
  }
  if (active && h == 0) {
    double acc[kAcc];
#pragma unroll
    for (int k = 0; k < kAcc; ++k) acc[k] = static_cast<double>(hi[k]) + static_cast<double>(lo[k]);
    Rate64 r;
    r.ux = kQ * acc[0];
    r.uy = kQ * acc[1];
    r.uz = kQ * acc[2];
    // PSE: V_t sum c a_s - a_t sum c V_s, the products exact in FP64
    const double tv = me.vol;
    r.wx = kQ * acc[3] + (tv * pse.x - static_cast<double>(me.ax) * pse.w);
    r.wy = kQ * acc[4] + (tv * pse.y - static_cast<double>(me.ay) * pse.w);
    r.wz = kQ * acc[5] + (tv * pse.z - static_cast<double>(me.az) * pse.w);
    if (it.slot < 0) a.out[t] = r;
    else a.partial[static_cast<size_t>(it.slot) * kLeafSize + self] = r;
  }
}

template <bool HighOrder, bool Uniform = false>
__device__ __forceinline__ void near_entry(const NearArgs& a, SrcRel* tile, float4* tab) {
  if (!HighOrder) {  // the core table into LDS (row halves; times ke for one radius, s_c kept)
    static_assert((2 * kCorePieces) % kNearSlice == 0, "whole passes over the table");
    for (int l = static_cast<int>(threadIdx.x); l < 2 * kCorePieces; l += kNearSlice) {
      const float* row = &kCoreTable[l >> 1][4 * (l & 1)];
      const float sc = Uniform ? a.u_ke : 1.0f;
      tab[l] = (l & 1) ? make_float4(row[0] * sc, Uniform ? a.u_lk3 : row[1], row[2], row[3])
                       : make_float4(row[0] * sc, row[1] * sc, row[2] * sc, row[3] * sc);
    }
    __syncthreads();
  }
  const int slices = kNearSlices * *a.total;
  for (;;) {
    int idx = 0;
    if (threadIdx.x == 0) idx = atomicAdd(a.queue, 1);
    idx = __builtin_amdgcn_readfirstlane(idx);  // lane 0's: every lane is active here
    if (idx >= slices) break;
    const NearItem it = a.items[idx / kNearSlices];
    const int slice = idx % kNearSlices;
    const int left = a.nodes[it.T].last - a.nodes[it.T].first - kNearSlice * slice;  // targets from here on
    if (left <= 0) continue;
    if (left <= kNearSlice / 2) near_slice<HighOrder, Uniform, true>(a, it, slice, tile, tab);
    else near_slice<HighOrder, Uniform, false>(a, it, slice, tile, tab);
  }
}

// the core table at LDS offset 0 (a row's second half is then in reach of
// one read2 without an address add), the source tile after it
struct NearLds {
  float4 tab[2 * kCorePieces];
  SrcRel tile[kNearSlice];
};
__global__ __launch_bounds__(kNearSlice) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_near_fp32c(NearArgs a) {
  __shared__ NearLds lds;
  near_entry<false>(a, lds.tile, lds.tab);
}
// every particle of one radius
__global__ __launch_bounds__(kNearSlice) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_near_fp32c_u(NearArgs a) {
  __shared__ NearLds lds;
  near_entry<false, true>(a, lds.tile, lds.tab);
}
__global__ __launch_bounds__(kNearSlice) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_near_fp32c_ho(NearArgs a) {
  __shared__ SrcRel tile[kNearSlice];
  near_entry<true>(a, tile, nullptr);
}

// rates of a split leaf: sum of its items' partial sums, in item order
__global__ __launch_bounds__(kTB) void k_near_reduce(const NodeRec* __restrict__ nodes, const int* __restrict__ parts,
                                                     const int* __restrict__ offs, const Rate64* __restrict__ partial,
                                                     Rate64* __restrict__ out) {
  const int T = blockIdx.x;
  const int P = parts[T];
  if (P <= 1) return;
  const int t = nodes[T].first + static_cast<int>(threadIdx.x);
  if (t >= nodes[T].last) return;
  const size_t base = static_cast<size_t>(offs[T]) * kTB + threadIdx.x;
  Rate64 s = partial[base];
  for (int p = 1; p < P; ++p) {
    const Rate64 q = partial[base + static_cast<size_t>(p) * kTB];
    s.ux += q.ux; s.uy += q.uy; s.uz += q.uz;
    s.wx += q.wx; s.wy += q.wy; s.wz += q.wz;
  }
  out[t] = s;
}

// ---------------------------------------------------------------------
// Entry kernels.  A grid kernel runs item blockIdx.x; a persistent kernel
// keeps its blocks resident and pulls items from an atomic queue (items
// are ordered largest first), so uneven items leave no idle tail.
// ---------------------------------------------------------------------
template <int Kind>
struct TileOf {
  using type = Rec64;
};
template <>
struct TileOf<kPairFP64> {
  using type = TileF64;
};
template <>
struct TileOf<kPairFP32C> {
  using type = SrcT2;
};
template <>
struct TileOf<kPairFP32CHO> {
  using type = SrcT2;
};
template <>
struct TileOf<kPairFP32CPlain> {
  using type = SrcT2;
};

template <int Kind>
__device__ __forceinline__ void run_item(const PairArgs& a, const WorkItem& it,
                                         typename TileOf<Kind>::type* tile) {
  if constexpr (Kind == kPairFP64Exact)
    item_fp64_ref(a, it, tile);
  else if constexpr (Kind == kPairFP64)
    item_fp64(a, it, tile);
  else if constexpr (Kind == kPairFP64HO)
    item_fp64_ho(a, it, tile);
  else if constexpr (Kind == kPairFP32CHO)
    item_fp32c<true>(a, it, tile);
  else if constexpr (Kind == kPairFP32CPlain)
    item_fp32c<false, false>(a, it, tile);
  else
    item_fp32c<false>(a, it, tile);
}

template <int Kind, bool Persistent>
__device__ __forceinline__ void pair_entry(const PairArgs& a) {
  __shared__ typename TileOf<Kind>::type tile[kTB];
  if (!Persistent) {
    run_item<Kind>(a, a.items[blockIdx.x], tile);
    return;
  }
  __shared__ int s_idx;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_idx = atomicAdd(a.queue, 1);
    __syncthreads();
    const int idx = s_idx;
    if (idx >= a.nitems) break;  // every lane sees the same idx: uniform exit
    run_item<Kind>(a, a.items[idx], tile);
  }
}

template <bool P>
__global__ __launch_bounds__(kTB) void k_pair_fp64_ref(PairArgs a) { pair_entry<kPairFP64Exact, P>(a); }
template <bool P>
__global__ __launch_bounds__(kTB) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_pair_fp64(PairArgs a) {
  pair_entry<kPairFP64, P>(a);
}
template <bool P>
__global__ __launch_bounds__(kTB) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_pair_fp64_ho(PairArgs a) {
  pair_entry<kPairFP64HO, P>(a);
}
template <bool P>
__global__ __launch_bounds__(64) void k_pair_fp64s(PairArgs a) { fp64s_entry<P>(a); }
// FP32C at the register budget of 7 waves/SIMD (no spills) and of 8
// waves/SIMD (one spill slot, touched once per tile)
template <bool P>
__global__ __launch_bounds__(kTB) void k_pair_fp32c(PairArgs a) { pair_entry<kPairFP32C, P>(a); }
template <bool P>
__global__ __launch_bounds__(kTB) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_pair_fp32c_w8(PairArgs a) {
  pair_entry<kPairFP32C, P>(a);
}
// the same two builds with the high-order algebraic kernel
template <bool P>
__global__ __launch_bounds__(kTB) void k_pair_fp32c_ho(PairArgs a) { pair_entry<kPairFP32CHO, P>(a); }
template <bool P>
__global__ __launch_bounds__(kTB) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_pair_fp32c_ho_w8(PairArgs a) {
  pair_entry<kPairFP32CHO, P>(a);
}

// FP32C fed through scalar loads: one-wave blocks, at the register budgets
// of 7 and 8 waves/SIMD, Gaussian and high-order kernels
template <bool P>
__global__ __launch_bounds__(64) void k_pair_fp32s(PairArgs a) { fp32s_entry<false, P>(a); }
template <bool P>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_pair_fp32s_w8(PairArgs a) {
  fp32s_entry<false, P>(a);
}
template <bool P>
__global__ __launch_bounds__(64) void k_pair_fp32s_ho(PairArgs a) { fp32s_entry<true, P>(a); }
// two targets per lane (every scalar source record feeds two pairs)
template <bool P>
__global__ __launch_bounds__(64) void k_pair_fp32s_t2(PairArgs a) { fp32s_entry<false, P, 2>(a); }
template <bool P>
__global__ __launch_bounds__(64) void k_pair_fp32s_ho_t2(PairArgs a) { fp32s_entry<true, P, 2>(a); }
template <bool P>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_pair_fp32s_ho_w8(PairArgs a) {
  fp32s_entry<true, P>(a);
}

// FP32 pairs summed without compensation (grid launch only): a measuring
// aid, not a production path
__global__ __launch_bounds__(kTB) void k_pair_fp32c_plain(PairArgs a) { pair_entry<kPairFP32CPlain, false>(a); }

// rates[t] = sum over parts of partial[p][t], parts in order (FP64).
__global__ __launch_bounds__(256) void k_reduce_parts(const Rate64* __restrict__ partial,
                                                      int parts, int n, Rate64* __restrict__ rates) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  Rate64 s = partial[t];
  for (int p = 1; p < parts; ++p) {
    const Rate64 q = partial[static_cast<size_t>(p) * n + t];
    s.ux += q.ux; s.uy += q.uy; s.uz += q.uz;
    s.wx += q.wx; s.wy += q.wy; s.wz += q.wz;
  }
  rates[t] = s;
}

}  // namespace

namespace {
// the scalar-fed FP32C launch: chunk records, then one wave per quarter item
hipError_t launch_fp32s(int kind, const PairConfig& cfg, const PairArgs& args, int grid, hipStream_t st) {
  if (!args.chunks || !args.src || args.nchunks <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_src_chunks, dim3(args.nchunks), dim3(64), 0, st, args.state, args.chunks, args.src);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const bool t2 = cfg.two_per_lane;
  const int waves = (t2 ? 2 : 4) * args.nitems;
  const bool P = cfg.persistent;
  if (P) {
    e = launch_fill_int(args.queue, 1, 0, st);
    if (e != hipSuccess) return e;
  }
  const dim3 g(P ? static_cast<unsigned>(std::min(grid, waves)) : static_cast<unsigned>(waves));
  const dim3 block(64);
  const bool ho = kind == kPairFP32CHO;
  if (t2 && ho) {
    if (P) hipLaunchKernelGGL(k_pair_fp32s_ho_t2<true>, g, block, 0, st, args);
    else hipLaunchKernelGGL(k_pair_fp32s_ho_t2<false>, g, block, 0, st, args);
  } else if (t2) {
    if (P) hipLaunchKernelGGL(k_pair_fp32s_t2<true>, g, block, 0, st, args);
    else hipLaunchKernelGGL(k_pair_fp32s_t2<false>, g, block, 0, st, args);
  } else if (ho && cfg.waves8) {
    if (P) hipLaunchKernelGGL(k_pair_fp32s_ho_w8<true>, g, block, 0, st, args);
    else hipLaunchKernelGGL(k_pair_fp32s_ho_w8<false>, g, block, 0, st, args);
  } else if (ho) {
    if (P) hipLaunchKernelGGL(k_pair_fp32s_ho<true>, g, block, 0, st, args);
    else hipLaunchKernelGGL(k_pair_fp32s_ho<false>, g, block, 0, st, args);
  } else if (cfg.waves8) {
    if (P) hipLaunchKernelGGL(k_pair_fp32s_w8<true>, g, block, 0, st, args);
    else hipLaunchKernelGGL(k_pair_fp32s_w8<false>, g, block, 0, st, args);
  } else {
    if (P) hipLaunchKernelGGL(k_pair_fp32s<true>, g, block, 0, st, args);
    else hipLaunchKernelGGL(k_pair_fp32s<false>, g, block, 0, st, args);
  }
  return hipGetLastError();
}
}  // namespace

hipError_t launch_pair(int kind, const PairConfig& cfg, const PairArgs& args, int grid,
                       hipStream_t st) {
  if (args.nitems <= 0) return hipSuccess;
  if (cfg.scalar_feed && kind == kPairFP64) {
    if (!args.chunks || !args.src64 || args.nchunks <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_src_chunks64, dim3(args.nchunks), dim3(64), 0, st, args.state, args.chunks, args.src64);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int waves = 4 * args.nitems;
    if (cfg.persistent) {
      e = launch_fill_int(args.queue, 1, 0, st);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL(k_pair_fp64s<true>, dim3(static_cast<unsigned>(std::min(grid, waves))), dim3(64), 0, st,
                         args);
    } else {
      hipLaunchKernelGGL(k_pair_fp64s<false>, dim3(static_cast<unsigned>(waves)), dim3(64), 0, st, args);
    }
    e = hipGetLastError();
    if (e != hipSuccess || cfg.parts <= 1) return e;
    hipLaunchKernelGGL(k_reduce_parts, dim3((args.n + 255) / 256), dim3(256), 0, st, args.partial,
                       cfg.parts, args.n, args.rates);
    return hipGetLastError();
  }
  if (cfg.scalar_feed && (kind == kPairFP32C || kind == kPairFP32CHO)) {
    hipError_t e = launch_fp32s(kind, cfg, args, grid, st);
    if (e != hipSuccess || cfg.parts <= 1) return e;
    hipLaunchKernelGGL(k_reduce_parts, dim3((args.n + 255) / 256), dim3(256), 0, st, args.partial,
                       cfg.parts, args.n, args.rates);
    return hipGetLastError();
  }
  const dim3 block(kTB);
  const dim3 g(cfg.persistent ? static_cast<unsigned>(std::min(grid, args.nitems))
                              : static_cast<unsigned>(args.nitems));
  if (cfg.persistent) {
    hipError_t e = launch_fill_int(args.queue, 1, 0, st);
    if (e != hipSuccess) return e;
  }
const bool P = cfg.persistent;
  switch (kind) {
    case kPairFP64Exact:
      if (P) hipLaunchKernelGGL(k_pair_fp64_ref<true>, g, block, 0, st, args);
      else hipLaunchKernelGGL(k_pair_fp64_ref<false>, g, block, 0, st, args);
      break;
    case kPairFP64:
      if (P) hipLaunchKernelGGL(k_pair_fp64<true>, g, block, 0, st, args);
      else hipLaunchKernelGGL(k_pair_fp64<false>, g, block, 0, st, args);
      break;
    case kPairFP64HO:
      if (P) hipLaunchKernelGGL(k_pair_fp64_ho<true>, g, block, 0, st, args);
      else hipLaunchKernelGGL(k_pair_fp64_ho<false>, g, block, 0, st, args);
      break;
    case kPairFP32CHO:
      if (cfg.waves8) {
        if (P) hipLaunchKernelGGL(k_pair_fp32c_ho_w8<true>, g, block, 0, st, args);
        else hipLaunchKernelGGL(k_pair_fp32c_ho_w8<false>, g, block, 0, st, args);
      } else {
        if (P) hipLaunchKernelGGL(k_pair_fp32c_ho<true>, g, block, 0, st, args);
        else hipLaunchKernelGGL(k_pair_fp32c_ho<false>, g, block, 0, st, args);
      }
      break;
    case kPairFP32CPlain:
      if (P || cfg.parts != 1) return hipErrorInvalidValue;
      hipLaunchKernelGGL(k_pair_fp32c_plain, g, block, 0, st, args);
      break;
    case kPairFP32C:
      if (cfg.waves8) {
        if (P) hipLaunchKernelGGL(k_pair_fp32c_w8<true>, g, block, 0, st, args);
        else hipLaunchKernelGGL(k_pair_fp32c_w8<false>, g, block, 0, st, args);
      } else {
        if (P) hipLaunchKernelGGL(k_pair_fp32c<true>, g, block, 0, st, args);
        else hipLaunchKernelGGL(k_pair_fp32c<false>, g, block, 0, st, args);
      }
      break;
    default:
      return hipErrorInvalidValue;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || cfg.parts <= 1) return e;
  hipLaunchKernelGGL(k_reduce_parts, dim3((args.n + 255) / 256), dim3(256), 0, st, args.partial,
                     cfg.parts, args.n, args.rates);
  return hipGetLastError();
}

hipError_t launch_point_velocity(const Rec64* state, int n, const double* pts, int m, double* u,
                                 bool high_order, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  if (high_order)
    hipLaunchKernelGGL(k_point_velocity<true>, dim3((m + kTB - 1) / kTB), dim3(kTB), 0, st, state, n, pts, m, u);
  else
    hipLaunchKernelGGL(k_point_velocity<false>, dim3((m + kTB - 1) / kTB), dim3(kTB), 0, st, state, n, pts, m, u);
  return hipGetLastError();
}

hipError_t launch_near_fp32c(const Rec64* sorted, const NodeRec* nodes, const TreeBuffers& tb, int nleaf,
                             double nu, float sigma_max2, bool uniform_sigma, bool highorder, int grid,
                             Rate64* rates_s, hipStream_t st) {
  if (nleaf <= 0) return hipSuccess;
  hipError_t e = launch_fill_int(tb.near_queue, 1, 0, st);
  if (e != hipSuccess) return e;
  NearArgs a{sorted, nodes, tb.near_list, tb.cap_near, tb.near_items, tb.near_total, tb.near_queue,
             static_cast<float>(nu), kFar32 * sigma_max2, rates_s, tb.near_partial, tb.leaf_chunks,
             0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  // near32u indexes the table without a clamp: every core sb must map below
  // kCorePieces (the largest core sb, in the device's FP32 roundings)
  const float sb_max = static_cast<float>(kFar32 * sigma_max2) * static_cast<float>(1.0 / sigma_max2);
  if (uniform_sigma && !highorder && sb_max * kCoreScale < static_cast<float>(kCorePieces)) {
    // the pair radius is sigma itself: sigma^2 = (s_s^2 + s_t^2) / 2; the
    // factors rounded from FP64, as near32t rounds them per pair from FP32
    const double s2 = sigma_max2, is2 = 1.0 / s2, is3 = is2 / std::sqrt(s2);
    a.u_cut2 = kFar32 * sigma_max2;
    const double ke = 0.79788456080286535588 * is3;  // sqrt(2/pi) / sigma^3
    a.u_isg2 = static_cast<float>(is2);
    a.u_ke = static_cast<float>(ke);
    a.u_lk3 = static_cast<float>(std::log2(ke / 3.0));
    a.u_ck3 = static_cast<float>(2.0 * nu * 0.063493635934240969785 * is3 * is2 / (ke / 3.0));  // (2 pi)^-3/2
    a.u_n3isg2 = static_cast<float>(-3.0 * is2);
    hipLaunchKernelGGL(k_near_fp32c_u, dim3(grid), dim3(kNearSlice), 0, st, a);
  } else if (highorder) {
    hipLaunchKernelGGL(k_near_fp32c_ho, dim3(grid), dim3(kNearSlice), 0, st, a);
  } else {
    hipLaunchKernelGGL(k_near_fp32c, dim3(grid), dim3(kNearSlice), 0, st, a);
  }
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_near_reduce, dim3(nleaf), dim3(kTB), 0, st, nodes, tb.near_parts, tb.near_offs,
                     tb.near_partial, rates_s);
  return hipGetLastError();
}

namespace {
// resident blocks per CU of a kernel (256 threads unless given), from the runtime's
// occupancy calculator (registers, LDS), cached per kernel
int blocks_per_cu(const void* kernel, int threads = kTB) {
  static std::map<const void*, int> cache;
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(kernel);
  if (it != cache.end()) return it->second;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, threads, 0) != hipSuccess || nb < 1)
    nb = kTB / threads * 4;
  cache[kernel] = nb;
  return nb;
}
}  // namespace

int pair_kernel_waves(int kind, const PairConfig& cfg) {
  const void* k = nullptr;
  if (cfg.scalar_feed && kind == kPairFP64) return blocks_per_cu(reinterpret_cast<const void*>(k_pair_fp64s<true>), 64);
  if (cfg.scalar_feed && (kind == kPairFP32C || kind == kPairFP32CHO)) {
    // one-wave blocks: resident blocks per CU = resident waves per CU
    if (cfg.two_per_lane)
      k = kind == kPairFP32C ? reinterpret_cast<const void*>(k_pair_fp32s_t2<true>)
                             : reinterpret_cast<const void*>(k_pair_fp32s_ho_t2<true>);
    else if (kind == kPairFP32C)
      k = cfg.waves8 ? reinterpret_cast<const void*>(k_pair_fp32s_w8<true>)
                     : reinterpret_cast<const void*>(k_pair_fp32s<true>);
    else
      k = cfg.waves8 ? reinterpret_cast<const void*>(k_pair_fp32s_ho_w8<true>)
                     : reinterpret_cast<const void*>(k_pair_fp32s_ho<true>);
    return blocks_per_cu(k, 64);
  }
  if (kind == kPairFP32C)
    k = cfg.waves8 ? reinterpret_cast<const void*>(k_pair_fp32c_w8<true>)
                   : reinterpret_cast<const void*>(k_pair_fp32c<true>);
  else if (kind == kPairFP32CHO)
    k = cfg.waves8 ? reinterpret_cast<const void*>(k_pair_fp32c_ho_w8<true>)
                   : reinterpret_cast<const void*>(k_pair_fp32c_ho<true>);
  else if (kind == kPairFP64)
    k = reinterpret_cast<const void*>(k_pair_fp64<true>);
  else if (kind == kPairFP64HO)
    k = reinterpret_cast<const void*>(k_pair_fp64_ho<true>);
  else
    k = reinterpret_cast<const void*>(k_pair_fp64_ref<true>);
  return blocks_per_cu(k);
}

int near_kernel_blocks_per_cu(bool highorder, bool uniform_sigma) {
  const void* k = highorder        ? reinterpret_cast<const void*>(k_near_fp32c_ho)
                  : uniform_sigma ? reinterpret_cast<const void*>(k_near_fp32c_u)
                                  : reinterpret_cast<const void*>(k_near_fp32c);
  return blocks_per_cu(k, kNearSlice);
}

}  // namespace pw
