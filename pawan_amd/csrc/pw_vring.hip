// pw_vring.hip -- strength resolution of a vortex ring on the device
// (__resolve::rebuild, src/resolve/resolve.cpp:12-63, on a __vring,
// src/wake/vring.cpp).
//
// A ring of m = (2L+1)^2 cross-section cells repeated over nAzi azimuthal
// stations carries one strength per cell: particle a m + i (station a,
// cell i) has vorticity s_i (-sin psi_a, cos psi_a, 0).  The reference
// builds the m x m influence matrix column by column: strengths e_i, the
// smoothed vorticity field of every particle (interaction.cpp:212-243), its
// in-plane magnitude at the first station's m particles (vring.cpp:208-215).
// Here one lane computes one entry M[p][i] directly: the field at particle p
// from cell i's nAzi particles in ascending station order (the terms of the
// other cells are exact zeros in the reference), with ZETASIG(0, s_p) for
// p itself.  The m x m system M s = ideal profile is then solved with
// partial pivoting (GSL's LU_decomp/LU_solve order) by one block for small
// rings and by every CU, one launch pair per column, for large ones, and
// the strengths are written into the ring's particles.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pw_internal.h"

namespace pw {
namespace {

constexpr int kTB = 256;
// matrices of this order and up are factored by all CUs, column by column
// (k_lu_pivot, k_lu_update), smaller ones by one block (k_lu_solve)
constexpr int kLuSpreadFrom = 256;
constexpr double k1Pi = 0.31830988618379067154;  // M_1_PI
constexpr double kP15 = 0x1.f7fccdff344acp+3;    // pow(2 pi, 1.5)

// ZETASIG of the Gaussian (gaussian.h:20-23) or of the high-order
// algebraic kernel (highorder.h:20-24, pow(x, 3.5) as x^3 sqrt(x)).
__device__ __forceinline__ double zetasig(double rho, double sigma, bool ho) {
  const double rb = rho / sigma;
  const double sig3 = sigma * sigma * sigma;
  if (ho) {
    const double b = rb * rb + 1.0;
    return 1.875 * k1Pi / (b * b * b * sqrt(b)) / sig3;
  }
  return exp(-0.5 * rb * rb) / sig3 / kP15;
}

// Smoothed vorticity field over work items (one block of up to 256 targets
// each, sources through LDS): INFLUENCE / SELFINFLUENCE
// (interaction_utils.h:626-703) from the target's side, with each pair's
// sigma = sqrt((s_s^2 + s_t^2)/2) and the target's own radius for itself.
__global__ __launch_bounds__(kTB) void k_vorticity_field(const Rec64* __restrict__ y,
                                                         const WorkItem* __restrict__ items, int ho,
                                                         double* __restrict__ field) {
  __shared__ Rec64 tile[kTB];
  const WorkItem it = items[blockIdx.x];
  const int t = it.t0 + static_cast<int>(threadIdx.x);
  const bool active = t < it.t1;
  const Rec64 me = y[active ? t : it.t0];
  double kx = 0.0, ky = 0.0, kz = 0.0;
  for (int base = it.s0; base < it.s1; base += kTB) {
    const int cnt = min(kTB, it.s1 - base);
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < cnt) tile[threadIdx.x] = y[base + threadIdx.x];
    __syncthreads();
    if (!active) continue;
    for (int l = 0; l < cnt; ++l) {
      const Rec64& s = tile[l];
      double Z;
      if (base + l == t) {
        Z = zetasig(0.0, sqrt(me.s2), ho != 0);
      } else {
        const double dx = me.x - s.x, dy = me.y - s.y, dz = me.z - s.z;
        const double rho = sqrt(dx * dx + dy * dy + dz * dz);
        Z = zetasig(rho, sqrt(0.5 * (s.s2 + me.s2)), ho != 0);
      }
      kx += Z * s.ax;
      ky += Z * s.ay;
      kz += Z * s.az;
    }
  }
  if (active) {
    field[3 * static_cast<size_t>(t)] += kx;
    field[3 * static_cast<size_t>(t) + 1] += ky;
    field[3 * static_cast<size_t>(t) + 2] += kz;
  }
}

// M[p * m + i] for p, i < m; ring particles start at `off` in the state.
__global__ __launch_bounds__(kTB) void k_vring_influence(const Rec64* __restrict__ y, int off, int m,
                                                         int nazi, double dpsi, int ho,
                                                         double* __restrict__ M) {
  const int e = blockIdx.x * kTB + static_cast<int>(threadIdx.x);
  if (e >= m * m) return;
  const int p = e / m, i = e - p * m;
  const Rec64 tp = y[off + p];
  double kx = 0.0, ky = 0.0;
  for (int a = 0; a < nazi; ++a) {
    const int q = a * m + i;
    const double psi = a * dpsi;
    const double tx = -sin(psi), ty = cos(psi);
    double Z;
    if (q == p) {
      Z = zetasig(0.0, sqrt(tp.s2), ho != 0);  // SELFINFLUENCE, interaction_utils.h:689-703
    } else {
      const Rec64 s = y[off + q];
      const double dx = s.x - tp.x, dy = s.y - tp.y, dz = s.z - tp.z;
      const double rho = sqrt(dx * dx + dy * dy + dz * dz);
      Z = zetasig(rho, sqrt(0.5 * (tp.s2 + s.s2)), ho != 0);
    }
    kx += Z * tx;
    ky += Z * ty;
  }
  M[e] = sqrt(kx * kx + ky * ky);
}

// Solve A x = b (n x n, row-major, overwritten by its LU factors) with one
// block: partial pivoting on the largest |a_ij| at or below the diagonal
// (lowest row on ties); each elimination step stores the multipliers in
// column j, then updates the trailing submatrix with the lanes along its
// rows (coalesced); the two triangular solves go column by column with the
// lanes over the rows, x held in LDS.  flag[0] = 1 if a pivot is zero or
// not finite (a NaN column keeps the diagonal row as its pivot, so no row
// outside the matrix is ever selected).
__global__ __launch_bounds__(kTB) void k_lu_solve(double* __restrict__ A, int n,
                                                  const double* __restrict__ b,
                                                  double* __restrict__ x, int* __restrict__ flag) {
#pragma clang fp contract(off)
  __shared__ int perm[kVringMaxCells];
  __shared__ double xs[kVringMaxCells];
  __shared__ double red_v[kTB];
  __shared__ int red_i[kTB];
  const int tid = static_cast<int>(threadIdx.x);
  for (int r = tid; r < n; r += kTB) perm[r] = r;
  if (tid == 0) flag[0] = 0;
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    double best = -1.0;
    int bi = j;
    for (int r = j + tid; r < n; r += kTB) {
      const double v = fabs(A[static_cast<size_t>(r) * n + j]);
      if (v > best) { best = v; bi = r; }
    }
    red_v[tid] = best;
    red_i[tid] = bi;
    __syncthreads();
    for (int s = kTB / 2; s > 0; s >>= 1) {
      if (tid < s) {
        const double v2 = red_v[tid + s];
        const int i2 = red_i[tid + s];
        if (v2 > red_v[tid] || (v2 == red_v[tid] && i2 < red_i[tid])) {
          red_v[tid] = v2;
          red_i[tid] = i2;
        }
      }
      __syncthreads();
    }
    const int ip = red_i[0];
    if (ip != j) {
      for (int c = tid; c < n; c += kTB) {
        const double t = A[static_cast<size_t>(j) * n + c];
        A[static_cast<size_t>(j) * n + c] = A[static_cast<size_t>(ip) * n + c];
        A[static_cast<size_t>(ip) * n + c] = t;
      }
      if (tid == 0) {
        const int t = perm[j];
        perm[j] = perm[ip];
        perm[ip] = t;
      }
    }
    __syncthreads();
    const double piv = A[static_cast<size_t>(j) * n + j];
    if (!(fabs(piv) > 0.0) || !isfinite(piv)) {
      if (tid == 0) flag[0] = 1;
      __syncthreads();
      continue;
    }
    for (int r = j + 1 + tid; r < n; r += kTB) A[static_cast<size_t>(r) * n + j] /= piv;
    __syncthreads();
    const int w = n - j - 1;
    const double* prow = A + static_cast<size_t>(j) * n;
    for (int e = tid; e < w * w; e += kTB) {
      const int r = j + 1 + e / w, c = j + 1 + e % w;
      double* row = A + static_cast<size_t>(r) * n;
      row[c] -= row[j] * prow[c];
    }
    __syncthreads();
  }
  for (int r = tid; r < n; r += kTB) xs[r] = b[perm[r]];
  __syncthreads();
  // L y = P b (unit diagonal), column by column
  for (int c = 0; c + 1 < n; ++c) {
    const double xc = xs[c];
    for (int r = c + 1 + tid; r < n; r += kTB) xs[r] -= A[static_cast<size_t>(r) * n + c] * xc;
    __syncthreads();
  }
  // U x = y, column by column from the last
  for (int c = n - 1; c >= 0; --c) {
    if (tid == 0) xs[c] /= A[static_cast<size_t>(c) * n + c];
    __syncthreads();
    const double xc = xs[c];
    for (int r = tid; r < c; r += kTB) xs[r] -= A[static_cast<size_t>(r) * n + c] * xc;
    __syncthreads();
  }
  for (int r = tid; r < n; r += kTB) x[r] = xs[r];
}

// The same factorization spread over the chip for larger matrices: per
// column j, k_lu_pivot (one block: pivot search, row swap, multipliers)
// and k_lu_update (every block: the trailing submatrix, one element per
// lane); then k_lu_substitute (one block: the two triangular solves).  Each
// element sees the same operations in the same order as in k_lu_solve, so
// the factors are bit for bit the same.  perm (n entries) lives in device
// memory between the launches.
__global__ __launch_bounds__(kTB) void k_lu_pivot(double* __restrict__ A, int n, int j, int* __restrict__ perm,
                                                  int* __restrict__ flag) {
#pragma clang fp contract(off)
  __shared__ double red_v[kTB];
  __shared__ int red_i[kTB];
  const int tid = static_cast<int>(threadIdx.x);
  if (j == 0) {
    for (int r = tid; r < n; r += kTB) perm[r] = r;
    if (tid == 0) flag[0] = 0;
    __syncthreads();
  }
  double best = -1.0;
  int bi = j;
  for (int r = j + tid; r < n; r += kTB) {
    const double v = fabs(A[static_cast<size_t>(r) * n + j]);
    if (v > best) { best = v; bi = r; }
  }
  red_v[tid] = best;
  red_i[tid] = bi;
  __syncthreads();
  for (int st = kTB / 2; st > 0; st >>= 1) {
    if (tid < st) {
      const double v2 = red_v[tid + st];
      const int i2 = red_i[tid + st];
      if (v2 > red_v[tid] || (v2 == red_v[tid] && i2 < red_i[tid])) {
        red_v[tid] = v2;
        red_i[tid] = i2;
      }
    }
    __syncthreads();
  }
  const int ip = red_i[0];
  if (ip != j) {
    for (int c = tid; c < n; c += kTB) {
      const double t = A[static_cast<size_t>(j) * n + c];
      A[static_cast<size_t>(j) * n + c] = A[static_cast<size_t>(ip) * n + c];
      A[static_cast<size_t>(ip) * n + c] = t;
    }
    if (tid == 0) {
      const int t = perm[j];
      perm[j] = perm[ip];
      perm[ip] = t;
    }
  }
  __syncthreads();
  const double piv = A[static_cast<size_t>(j) * n + j];
  if (!(fabs(piv) > 0.0) || !isfinite(piv)) {
    if (tid == 0) flag[0] = 1;
    return;
  }
  for (int r = j + 1 + tid; r < n; r += kTB) A[static_cast<size_t>(r) * n + j] /= piv;
}

__global__ __launch_bounds__(kTB) void k_lu_update(double* __restrict__ A, int n, int j) {
#pragma clang fp contract(off)
  const double piv = A[static_cast<size_t>(j) * n + j];
  if (!(fabs(piv) > 0.0) || !isfinite(piv)) return;  // k_lu_pivot flagged it and left the column
  const int w = n - j - 1;
  const double* prow = A + static_cast<size_t>(j) * n;
  for (int e = blockIdx.x * kTB + static_cast<int>(threadIdx.x); e < w * w; e += gridDim.x * kTB) {
    const int r = j + 1 + e / w, c = j + 1 + e % w;
    double* row = A + static_cast<size_t>(r) * n;
    row[c] -= row[j] * prow[c];
  }
}

__global__ __launch_bounds__(kTB) void k_lu_substitute(const double* __restrict__ A, int n,
                                                       const int* __restrict__ perm,
                                                       const double* __restrict__ b, double* __restrict__ x) {
#pragma clang fp contract(off)
  __shared__ double xs[kVringMaxCells];
  const int tid = static_cast<int>(threadIdx.x);
  for (int r = tid; r < n; r += kTB) xs[r] = b[perm[r]];
  __syncthreads();
  for (int c = 0; c + 1 < n; ++c) {
    const double xc = xs[c];
    for (int r = c + 1 + tid; r < n; r += kTB) xs[r] -= A[static_cast<size_t>(r) * n + c] * xc;
    __syncthreads();
  }
  for (int c = n - 1; c >= 0; --c) {
    if (tid == 0) xs[c] /= A[static_cast<size_t>(c) * n + c];
    __syncthreads();
    const double xc = xs[c];
    for (int r = tid; r < c; r += kTB) xs[r] -= A[static_cast<size_t>(r) * n + c] * xc;
    __syncthreads();
  }
  for (int r = tid; r < n; r += kTB) x[r] = xs[r];
}

// a = s_i (-sin psi_a, cos psi_a, 0) for the ring's nazi * m particles
// (__vring::set_vorticity, vring.cpp:25-48); left unchanged if the solve
// found the matrix singular.
__global__ __launch_bounds__(kTB) void k_vring_set_alpha(Rec64* __restrict__ y, int off, int m, int nazi,
                                                         double dpsi, const double* __restrict__ s,
                                                         const int* __restrict__ flag) {
  const int q = blockIdx.x * kTB + static_cast<int>(threadIdx.x);
  if (q >= m * nazi || flag[0] != 0) return;
  const int a = q / m, i = q - a * m;
  const double psi = a * dpsi;
  Rec64& r = y[off + q];
  r.ax = -sin(psi) * s[i];
  r.ay = cos(psi) * s[i];
  r.az = 0.0;
}

}  // namespace

hipError_t launch_vorticity_field(const Rec64* y, const WorkItem* items, int nitems, bool high_order,
                                  double* field, hipStream_t st) {
  if (nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_vorticity_field, dim3(nitems), dim3(kTB), 0, st, y, items, high_order ? 1 : 0, field);
  return hipGetLastError();
}

hipError_t launch_vring_resolve(Rec64* y, int off, int m, int nazi, bool high_order, double* M,
                                double* lu, const double* ideal, double* s, int* flag, hipStream_t st) {
  if (m <= 0 || m > kVringMaxCells || nazi <= 0) return hipErrorInvalidValue;
  const double dpsi = 2.0 * 3.14159265358979323846 / nazi;
  hipLaunchKernelGGL(k_vring_influence, dim3((m * m + kTB - 1) / kTB), dim3(kTB), 0, st, y, off, m, nazi,
                     dpsi, high_order ? 1 : 0, M);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(lu, M, sizeof(double) * m * m, hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return e;
  if (m < kLuSpreadFrom) {
    hipLaunchKernelGGL(k_lu_solve, dim3(1), dim3(kTB), 0, st, lu, m, ideal, s, flag);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  } else {
    int* perm = flag + 1;
    for (int j = 0; j < m; ++j) {
      hipLaunchKernelGGL(k_lu_pivot, dim3(1), dim3(kTB), 0, st, lu, m, j, perm, flag);
      const int w = m - j - 1;
      if (w > 0) {
        const int blocks = std::min((w * w + kTB - 1) / kTB, 1024);
        hipLaunchKernelGGL(k_lu_update, dim3(blocks), dim3(kTB), 0, st, lu, m, j);
      }
    }
    hipLaunchKernelGGL(k_lu_substitute, dim3(1), dim3(kTB), 0, st, lu, m, perm, ideal, s);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_vring_set_alpha, dim3((m * nazi + kTB - 1) / kTB), dim3(kTB), 0, st, y, off, m, nazi,
                     dpsi, s, flag);
  return hipGetLastError();
}

}  // namespace pw
