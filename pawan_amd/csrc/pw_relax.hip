// pw_relax.hip -- particle management of one wake on the device: split and
// merge (__wake::split / merge / relax, src/wake/wake.cpp:61-165), behind
// pw_wake_split and pw_wake_merge (include/pawan_gpu.h).
//
// Split.  Whether a particle splits depends on that particle alone, so every
// particle of the age prefix is tested at once; an exclusive scan of the
// decisions gives each new particle its row (new particles in the order of
// their parents, after the last particle, as the reference appends them).
//
// Merge.  The reference walks particle i over all later j, moving x_i and
// growing a_i as it merges, so later tests see the moved x_i: the result
// depends on the order.  It is reproduced exactly in two steps.
//  1. A cell list (cells of the largest merge distance, (sigma_i +
//     sigma_j)/8 <= sigma_max/4, keyed by cell and sorted) lets every i of the
//     age prefix test its later neighbours in parallel with the original
//     state.  The first merge of i needs such a neighbour (x_i is unmoved
//     until then, and strengths only ever drop to zero), so an i with none
//     does nothing and is left out.
//  2. One block then visits the flagged i in order and replays the
//     reference's inner loop: the next j > (last j seen) within the merge
//     distance of the current x_i, found among the 27 cells around x_i, with
//     the current strengths.  Later particles have not moved (only the outer
//     particle of an iteration moves), so the cell list of the original
//     positions still finds them.
// The arithmetic is the reference's (GSL's dnrm2 and ddot, dscal by the
// reciprocal, no contraction), so the result is bit-for-bit the
// sequential one.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "pw_ctx.h"

namespace {

using pwi::fail;

constexpr int kB = 256;
constexpr int kCellBits = 21;  // cell coordinate bits per axis of a key
constexpr int64_t kCellMax = (int64_t(1) << kCellBits) - 1;
constexpr uint64_t kNoCell = ~uint64_t(0);

#pragma clang fp contract(off)

// gsl_blas_dnrm2 of a 3-vector: the netlib scaled sum of squares
__device__ double dnrm2_3(double x0, double x1, double x2) {
  double scale = 0.0, ssq = 1.0;
  const double x[3] = {x0, x1, x2};
  for (int i = 0; i < 3; ++i) {
    if (x[i] != 0.0) {
      const double ax = fabs(x[i]);
      if (scale < ax) {
        ssq = 1.0 + ssq * (scale / ax) * (scale / ax);
        scale = ax;
      } else {
        ssq += (ax / scale) * (ax / scale);
      }
    }
  }
  return scale * sqrt(ssq);
}

// the age test of both loops (wake.cpp:64, 115): stepnum - active > 45
__global__ void k_age_prefix(const double* __restrict__ active, int n, double stepnum, int* __restrict__ prefix) {
  const int i = blockIdx.x * kB + static_cast<int>(threadIdx.x);
  if (i < n && !(stepnum - active[i] > 45)) atomicMin(prefix, i);
}

// ---- split ----------------------------------------------------------------

__global__ void k_split_flags(int n, const int* __restrict__ prefix, const double* __restrict__ alpha,
                              const double* __restrict__ birth, int* __restrict__ flag) {
  const int i = blockIdx.x * kB + static_cast<int>(threadIdx.x);
  if (i >= n) return;
  int f = 0;
  if (i < *prefix) {
    const double ib = birth[i];
    const double m = dnrm2_3(alpha[3 * i], alpha[3 * i + 1], alpha[3 * i + 2]);
    f = (m >= 2 * ib && ib > 0.0001) ? 1 : 0;
  }
  flag[i] = f;
}

// parent i (flagged) and its new particle at row first + off[i]
__global__ void k_split_apply(int n, int64_t first, const int* __restrict__ flag, const int* __restrict__ off,
                              double* __restrict__ pos, double* __restrict__ alpha, double* __restrict__ radius,
                              double* __restrict__ volume, double* __restrict__ birth, double* __restrict__ active) {
  const int i = blockIdx.x * kB + static_cast<int>(threadIdx.x);
  if (i >= n || !flag[i]) return;
  const int64_t k = first + off[i];
  double* ip = pos + 3 * static_cast<int64_t>(i);
  double* iv = alpha + 3 * static_cast<int64_t>(i);
  const double im = dnrm2_3(iv[0], iv[1], iv[2]);
  const double ir = radius[i];
  const double sc = 0.25 * ir / im;  // gsl_blas_dscal(0.25*iradius/ivormag, deltapos)
  double dp[3], x[3], a[3];
  for (int c = 0; c < 3; ++c) {
    dp[c] = iv[c] * sc;
    x[c] = ip[c];
    a[c] = iv[c];
  }
  for (int c = 0; c < 3; ++c) {
    pos[3 * k + c] = x[c] + dp[c];
    alpha[3 * k + c] = a[c] * 0.5;
    ip[c] = x[c] - dp[c];
    iv[c] = a[c] * 0.5;
  }
  radius[k] = ir;
  volume[k] = volume[i];
  birth[k] = 0.5 * im;
  active[k] = active[i];
  birth[i] = 0.5 * im;
}

// ---- merge ----------------------------------------------------------------

// bounding box and largest radius, one block (n is at most 2^31)
__global__ __launch_bounds__(1024) void k_extent(const double* __restrict__ pos, const double* __restrict__ radius,
                                                 int n, double* __restrict__ out) {
  __shared__ double red[7][1024];
  double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
  for (int i = static_cast<int>(threadIdx.x); i < n; i += 1024) {
    for (int c = 0; c < 3; ++c) {
      const double x = pos[3 * static_cast<int64_t>(i) + c];
      if (isfinite(x)) {
        v[c] = fmin(v[c], x);
        v[3 + c] = fmax(v[3 + c], x);
      }
    }
    const double r = radius[i];
    if (isfinite(r)) v[6] = fmax(v[6], r);
  }
  for (int k = 0; k < 7; ++k) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s)
      for (int k = 0; k < 7; ++k)
        red[k][threadIdx.x] = k < 3 ? fmin(red[k][threadIdx.x], red[k][threadIdx.x + s])
                                    : fmax(red[k][threadIdx.x], red[k][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x < 7) out[threadIdx.x] = red[threadIdx.x][0];
}

struct Grid {
  double x0, y0, z0, h;  // origin and cell size
};

__device__ __forceinline__ uint64_t cell_key(int64_t ix, int64_t iy, int64_t iz) {
  return (static_cast<uint64_t>(iz) << (2 * kCellBits)) | (static_cast<uint64_t>(iy) << kCellBits) |
         static_cast<uint64_t>(ix);
}

__device__ __forceinline__ bool cell_of(const Grid& g, double x, double y, double z, int64_t* c) {
  if (!(isfinite(x) && isfinite(y) && isfinite(z))) return false;
  c[0] = static_cast<int64_t>(floor((x - g.x0) / g.h));
  c[1] = static_cast<int64_t>(floor((y - g.y0) / g.h));
  c[2] = static_cast<int64_t>(floor((z - g.z0) / g.h));
  for (int k = 0; k < 3; ++k) c[k] = min(max(c[k], int64_t(0)), kCellMax);
  return true;
}

__global__ void k_cell_keys(const double* __restrict__ pos, int n, Grid g, uint64_t* __restrict__ key,
                            int* __restrict__ idx) {
  const int i = blockIdx.x * kB + static_cast<int>(threadIdx.x);
  if (i >= n) return;
  int64_t c[3];
  const double* x = pos + 3 * static_cast<int64_t>(i);
  key[i] = cell_of(g, x[0], x[1], x[2], c) ? cell_key(c[0], c[1], c[2]) : kNoCell;
  idx[i] = i;
}

// [lo, hi) of the sorted keys equal to k
__device__ __forceinline__ void equal_range(const uint64_t* __restrict__ keys, int n, uint64_t k, int* lo, int* hi) {
  int a = 0, b = n;
  while (a < b) {
    const int m = (a + b) >> 1;
    if (keys[m] < k) a = m + 1;
    else b = m;
  }
  int c = a, d = n;
  while (c < d) {
    const int m = (c + d) >> 1;
    if (keys[m] <= k) c = m + 1;
    else d = m;
  }
  *lo = a;
  *hi = c;
}

// can pair (i, j) merge (wake.cpp:129-135), for positions xi, xj and
// strengths ai, aj with norms mi, mj
__device__ __forceinline__ bool in_reach(const double* xi, const double* xj, double ri, double rj) {
  const double rho = dnrm2_3(xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]);
  return rho < 0.125 * (ri + rj);
}
__device__ __forceinline__ double misalignment(const double* ai, const double* aj, double mi, double mj) {
  double al = 0.0;  // gsl_blas_ddot
  al += ai[0] * aj[0];
  al += ai[1] * aj[1];
  al += ai[2] * aj[2];
  return 1 - al / mi / mj;
}

// step 1: particles of the age prefix that have a later neighbour they
// would merge with in the original state
__global__ void k_merge_flags(int n, const int* __restrict__ prefix, Grid g, const uint64_t* __restrict__ keys,
                              const int* __restrict__ sidx, const double* __restrict__ pos,
                              const double* __restrict__ alpha, const double* __restrict__ radius,
                              int* __restrict__ flag) {
  const int i = blockIdx.x * kB + static_cast<int>(threadIdx.x);
  if (i >= n) return;
  int f = 0;
  const double* xi = pos + 3 * static_cast<int64_t>(i);
  const double* ai = alpha + 3 * static_cast<int64_t>(i);
  const double mi = dnrm2_3(ai[0], ai[1], ai[2]);
  int64_t c[3];
  if (i < *prefix && mi != 0.0 && cell_of(g, xi[0], xi[1], xi[2], c)) {
    const double ri = radius[i];
    for (int dz = -1; dz <= 1 && !f; ++dz)
      for (int dy = -1; dy <= 1 && !f; ++dy)
        for (int dx = -1; dx <= 1 && !f; ++dx) {
          const int64_t x = c[0] + dx, y = c[1] + dy, z = c[2] + dz;
          if (x < 0 || y < 0 || z < 0 || x > kCellMax || y > kCellMax || z > kCellMax) continue;
          int lo, hi;
          equal_range(keys, n, cell_key(x, y, z), &lo, &hi);
          for (int e = lo; e < hi && !f; ++e) {
            const int j = sidx[e];
            if (j <= i) continue;
            const double* aj = alpha + 3 * static_cast<int64_t>(j);
            const double mj = dnrm2_3(aj[0], aj[1], aj[2]);
            if (mj == 0.0 || !in_reach(xi, pos + 3 * static_cast<int64_t>(j), ri, radius[j])) continue;
            if (misalignment(ai, aj, mi, mj) >= 0.0001) f = 1;
          }
        }
  }
  flag[i] = f;
}

constexpr int kRT = 256;    // threads of the replaying block
constexpr int kCap = 1024;  // later neighbours of one outer particle kept in LDS

// the smallest of the block's values (all threads call it)
__device__ __forceinline__ int block_min_int(int v, int* red) {
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  int m = red[0];
  for (int w = 1; w < kRT / 64; ++w) m = min(m, red[w]);
  __syncthreads();
  return m;
}

// step 2: one block replays the reference's loops over the flagged
// particles.  For the current outer particle the later particles of the 27
// cells around it (index, position, radius, |a| at that time) are gathered
// into LDS, again when x_i moves to another cell; each step then takes the
// smallest j past the last one seen that is within reach of the current x_i.
// More than kCap of them: the step scans the cells in memory instead.
__global__ __launch_bounds__(kRT) void k_merge_resolve(int n, const int* __restrict__ list,
                                                       const int* __restrict__ count, Grid g,
                                                       const uint64_t* __restrict__ keys,
                                                       const int* __restrict__ sidx, double* pos, double* alpha,
                                                       const double* __restrict__ radius, double* volume,
                                                       int64_t* __restrict__ merges) {
  __shared__ int s_lo[27], s_hi[27];
  __shared__ int s_cnt;
  __shared__ int s_j[kCap];
  __shared__ double s_x[3 * kCap];
  __shared__ double s_r[kCap];
  __shared__ int s_red[kRT / 64];
  const int tid = static_cast<int>(threadIdx.x);
  const int K = *count;
  int64_t done = 0;
  for (int k = 0; k < K; ++k) {
    const int i = list[k];
    double xi[3], ai[3];
    for (int c = 0; c < 3; ++c) {
      xi[c] = pos[3 * static_cast<int64_t>(i) + c];
      ai[c] = alpha[3 * static_cast<int64_t>(i) + c];
    }
    const double mi = dnrm2_3(ai[0], ai[1], ai[2]);  // once per i, as the reference takes it
    if (mi == 0.0) continue;
    const double ri = radius[i];
    int jlast = i;
    int64_t cur[3] = {-1, -1, -1};
    bool have = false;
    for (;;) {
      int64_t c[3];
      const bool in_grid = cell_of(g, xi[0], xi[1], xi[2], c);
      if (!in_grid) break;  // a non-finite position reaches nothing
      if (!have || c[0] != cur[0] || c[1] != cur[1] || c[2] != cur[2]) {
        // the cell ranges, then the later particles of those cells
        if (tid < 27) {
          const int64_t x = c[0] + tid % 3 - 1, y = c[1] + (tid / 3) % 3 - 1, z = c[2] + tid / 9 - 1;
          if (x < 0 || y < 0 || z < 0 || x > kCellMax || y > kCellMax || z > kCellMax) {
            s_lo[tid] = s_hi[tid] = 0;
          } else {
            equal_range(keys, n, cell_key(x, y, z), &s_lo[tid], &s_hi[tid]);
          }
        }
        if (tid == 0) s_cnt = 0;
        __syncthreads();
        for (int q = 0; q < 27; ++q)
          for (int e = s_lo[q] + tid; e < s_hi[q]; e += kRT) {
            const int j = sidx[e];
            if (j <= jlast) continue;
            const int slot = atomicAdd(&s_cnt, 1);
            if (slot < kCap) {
              s_j[slot] = j;
              for (int c2 = 0; c2 < 3; ++c2) s_x[3 * slot + c2] = pos[3 * static_cast<int64_t>(j) + c2];
              s_r[slot] = radius[j];
            }
          }
        __syncthreads();
        for (int c2 = 0; c2 < 3; ++c2) cur[c2] = c[c2];
        have = true;
      }
      // the smallest j > jlast within reach of the current x_i, with a
      // non-zero strength
      int best = INT_MAX;
      const int cnt = s_cnt;
      if (cnt <= kCap) {
        for (int e = tid; e < cnt; e += kRT) {
          const int j = s_j[e];
          if (j <= jlast || j >= best) continue;
          if (!in_reach(xi, &s_x[3 * e], ri, s_r[e])) continue;
          const double* aj = alpha + 3 * static_cast<int64_t>(j);
          if (dnrm2_3(aj[0], aj[1], aj[2]) != 0.0) best = j;
        }
      } else {
        for (int q = 0; q < 27; ++q)
          for (int e = s_lo[q] + tid; e < s_hi[q]; e += kRT) {
            const int j = sidx[e];
            if (j <= jlast || j >= best) continue;
            if (!in_reach(xi, pos + 3 * static_cast<int64_t>(j), ri, radius[j])) continue;
            const double* aj = alpha + 3 * static_cast<int64_t>(j);
            if (dnrm2_3(aj[0], aj[1], aj[2]) != 0.0) best = j;
          }
      }
      const int j = block_min_int(best, s_red);
      if (j == INT_MAX) break;
      double xj[3], aj[3];
      for (int c2 = 0; c2 < 3; ++c2) {
        xj[c2] = pos[3 * static_cast<int64_t>(j) + c2];
        aj[c2] = alpha[3 * static_cast<int64_t>(j) + c2];
      }
      const double mj = dnrm2_3(aj[0], aj[1], aj[2]);
      if (misalignment(ai, aj, mi, mj) >= 0.0001) {
        const double sc = 1 / (mi + mj);
        for (int c2 = 0; c2 < 3; ++c2) {
          xi[c2] = (xi[c2] * mi + xj[c2] * mj) * sc;
          ai[c2] = ai[c2] + aj[c2];
        }
        __syncthreads();  // every thread has read a_j
        if (tid == 0) {
          for (int c2 = 0; c2 < 3; ++c2) alpha[3 * static_cast<int64_t>(j) + c2] = 0.0;
          volume[j] = 0.0;
        }
        __threadfence_block();
        __syncthreads();
        ++done;
      }
      jlast = j;
    }
    __syncthreads();
    if (tid == 0)
      for (int c2 = 0; c2 < 3; ++c2) {
        pos[3 * static_cast<int64_t>(i) + c2] = xi[c2];
        alpha[3 * static_cast<int64_t>(i) + c2] = ai[c2];
      }
    __threadfence_block();
    __syncthreads();
  }
  if (tid == 0) *merges = done;
}

#pragma clang fp contract(on)

// device buffers of one call, freed on every path
struct Buffers {
  std::vector<void*> p;
  hipStream_t st = nullptr;
  template <typename T>
  bool get(T** out, size_t count) {
    *out = nullptr;
    if (pwi::dalloc(out, count) != hipSuccess) return false;
    p.push_back(*out);
    return true;
  }
  ~Buffers() {
    if (st) (void)hipStreamSynchronize(st);
    for (void* q : p) (void)hipFree(q);
    if (st) (void)hipStreamDestroy(st);
  }
};

int open_device(int device, Buffers& b) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(PW_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(PW_ERR_ARG, "device index out of range");
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b.st, hipStreamNonBlocking) != hipSuccess)
    return fail(PW_ERR_HIP, "stream creation failed");
  return PW_OK;
}

#define RX_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) return fail(PW_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
hipError_t up(T* d, const T* h, size_t count, hipStream_t st) {
  return hipMemcpyAsync(d, h, count * sizeof(T), hipMemcpyHostToDevice, st);
}
template <typename T>
hipError_t down(T* h, const T* d, size_t count, hipStream_t st) {
  return hipMemcpyAsync(h, d, count * sizeof(T), hipMemcpyDeviceToHost, st);
}

}  // namespace

extern "C" {

int pw_wake_split(int device, double stepnum, int64_t maxn, int64_t* num_particles, int64_t* npidx, double* pos,
                  double* alpha, double* radius, double* volume, double* birthstrength, double* active,
                  int64_t* nsplit) {
  if (!num_particles || !npidx || !pos || !alpha || !radius || !volume || !birthstrength || !active)
    return fail(PW_ERR_ARG, "null argument");
  const int64_t n = *num_particles;
  if (n < 0 || maxn < n || maxn > INT32_MAX / 2) return fail(PW_ERR_ARG, "need 0 <= num_particles <= maxn < 2^30");
  if (*npidx != n)
    return fail(PW_ERR_UNSUPPORTED, "new particles are appended after the last one: npidx must equal "
                                    "num_particles (the reference writes them at npidx, wake.cpp:81-99)");
  if (nsplit) *nsplit = 0;
  if (n == 0) return PW_OK;
  Buffers b;
  PW_TRY(open_device(device, b));
  const size_t m = static_cast<size_t>(maxn);
  double *d_pos, *d_alpha, *d_rad, *d_vol, *d_birth, *d_act;
  int *d_prefix, *d_flag, *d_off;
  if (!b.get(&d_pos, 3 * m) || !b.get(&d_alpha, 3 * m) || !b.get(&d_rad, m) || !b.get(&d_vol, m) ||
      !b.get(&d_birth, m) || !b.get(&d_act, m) || !b.get(&d_prefix, 1) || !b.get(&d_flag, n) || !b.get(&d_off, n)) {
    (void)hipGetLastError();
    return fail(PW_ERR_NOMEM, "device buffers");
  }
  const size_t nn = static_cast<size_t>(n);
  RX_HIP(up(d_pos, pos, 3 * nn, b.st));
  RX_HIP(up(d_alpha, alpha, 3 * nn, b.st));
  RX_HIP(up(d_rad, radius, nn, b.st));
  RX_HIP(up(d_vol, volume, nn, b.st));
  RX_HIP(up(d_birth, birthstrength, nn, b.st));
  RX_HIP(up(d_act, active, nn, b.st));
  const int ni = static_cast<int>(n);
  const dim3 grid((ni + kB - 1) / kB);
  RX_HIP(pw::launch_fill_int(d_prefix, 1, ni, b.st));
  hipLaunchKernelGGL(k_age_prefix, grid, dim3(kB), 0, b.st, d_act, ni, stepnum, d_prefix);
  RX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_split_flags, grid, dim3(kB), 0, b.st, ni, d_prefix, d_alpha, d_birth, d_flag);
  RX_HIP(hipGetLastError());
  size_t tb = 0;
  RX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, d_flag, d_off, ni, b.st));
  void* d_temp = nullptr;
  if (!b.get(reinterpret_cast<char**>(&d_temp), tb)) return fail(PW_ERR_NOMEM, "scan buffer");
  RX_HIP(hipcub::DeviceScan::ExclusiveSum(d_temp, tb, d_flag, d_off, ni, b.st));
  int last_off = 0, last_flag = 0;
  RX_HIP(down(&last_off, d_off + (ni - 1), 1, b.st));
  RX_HIP(down(&last_flag, d_flag + (ni - 1), 1, b.st));
  RX_HIP(hipStreamSynchronize(b.st));
  const int64_t s = static_cast<int64_t>(last_off) + last_flag;
  if (n + s > maxn)
    return fail(PW_ERR_UNSUPPORTED, "the new particles would not fit in maxn rows (the reference then wraps "
                                    "and overwrites its first particles, wake.cpp:96-99)");
  if (s > 0) {
    hipLaunchKernelGGL(k_split_apply, grid, dim3(kB), 0, b.st, ni, n, d_flag, d_off, d_pos, d_alpha, d_rad, d_vol,
                       d_birth, d_act);
    RX_HIP(hipGetLastError());
    const size_t nt = static_cast<size_t>(n + s);
    RX_HIP(down(pos, d_pos, 3 * nt, b.st));
    RX_HIP(down(alpha, d_alpha, 3 * nt, b.st));
    RX_HIP(down(radius, d_rad, nt, b.st));
    RX_HIP(down(volume, d_vol, nt, b.st));
    RX_HIP(down(birthstrength, d_birth, nt, b.st));
    RX_HIP(down(active, d_act, nt, b.st));
    RX_HIP(hipStreamSynchronize(b.st));
  }
  *num_particles = n + s;
  *npidx = n + s;
  if (nsplit) *nsplit = s;
  return PW_OK;
}

int pw_wake_merge(int device, double stepnum, int64_t num_particles, double* pos, double* alpha,
                  const double* radius, double* volume, const double* active, int64_t* nmerged) {
  if (!pos || !alpha || !radius || !volume || !active) return fail(PW_ERR_ARG, "null argument");
  if (num_particles < 0 || num_particles > INT32_MAX / 2) return fail(PW_ERR_ARG, "need 0 <= num_particles < 2^30");
  if (nmerged) *nmerged = 0;
  const int64_t n = num_particles;
  if (n < 2) return PW_OK;
  Buffers b;
  PW_TRY(open_device(device, b));
  const size_t nn = static_cast<size_t>(n);
  const int ni = static_cast<int>(n);
  double *d_pos, *d_alpha, *d_rad, *d_vol, *d_act, *d_ext;
  int *d_prefix, *d_flag, *d_idx, *d_sidx, *d_list, *d_count;
  uint64_t *d_key, *d_skey;
  int64_t* d_merges;
  if (!b.get(&d_pos, 3 * nn) || !b.get(&d_alpha, 3 * nn) || !b.get(&d_rad, nn) || !b.get(&d_vol, nn) ||
      !b.get(&d_act, nn) || !b.get(&d_ext, 7) || !b.get(&d_prefix, 1) || !b.get(&d_flag, nn) ||
      !b.get(&d_idx, nn) || !b.get(&d_sidx, nn) || !b.get(&d_list, nn) || !b.get(&d_count, 1) ||
      !b.get(&d_key, nn) || !b.get(&d_skey, nn) || !b.get(&d_merges, 1)) {
    (void)hipGetLastError();
    return fail(PW_ERR_NOMEM, "device buffers");
  }
  RX_HIP(up(d_pos, pos, 3 * nn, b.st));
  RX_HIP(up(d_alpha, alpha, 3 * nn, b.st));
  RX_HIP(up(d_rad, radius, nn, b.st));
  RX_HIP(up(d_vol, volume, nn, b.st));
  RX_HIP(up(d_act, active, nn, b.st));
  // the cell size: the largest merge distance, with room for rounding
  hipLaunchKernelGGL(k_extent, dim3(1), dim3(1024), 0, b.st, d_pos, d_rad, ni, d_ext);
  RX_HIP(hipGetLastError());
  double ext[7];
  RX_HIP(down(ext, d_ext, 7, b.st));
  RX_HIP(hipStreamSynchronize(b.st));
  if (!(ext[6] > 0.0)) return PW_OK;  // no radius: nothing is within reach
  Grid g;
  g.x0 = std::isfinite(ext[0]) ? ext[0] : 0.0;
  g.y0 = std::isfinite(ext[1]) ? ext[1] : 0.0;
  g.z0 = std::isfinite(ext[2]) ? ext[2] : 0.0;
  double span = 0.0;
  for (int c = 0; c < 3; ++c)
    if (std::isfinite(ext[c]) && std::isfinite(ext[3 + c])) span = std::max(span, ext[3 + c] - ext[c]);
  g.h = std::max(0.25 * ext[6] * (1.0 + 1e-6), span / static_cast<double>(kCellMax - 1));
  const dim3 grid((ni + kB - 1) / kB);
  RX_HIP(pw::launch_fill_int(d_prefix, 1, ni, b.st));
  hipLaunchKernelGGL(k_age_prefix, grid, dim3(kB), 0, b.st, d_act, ni, stepnum, d_prefix);
  RX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_cell_keys, grid, dim3(kB), 0, b.st, d_pos, ni, g, d_key, d_idx);
  RX_HIP(hipGetLastError());
  size_t t1 = 0, t2 = 0;
  RX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, d_key, d_skey, d_idx, d_sidx, ni, 0, 64, b.st));
  RX_HIP(hipcub::DeviceSelect::Flagged(nullptr, t2, d_idx, d_flag, d_list, d_count, ni, b.st));
  char* d_temp = nullptr;
  if (!b.get(&d_temp, std::max(t1, t2))) return fail(PW_ERR_NOMEM, "sort buffer");
  RX_HIP(hipcub::DeviceRadixSort::SortPairs(d_temp, t1, d_key, d_skey, d_idx, d_sidx, ni, 0, 64, b.st));
  hipLaunchKernelGGL(k_merge_flags, grid, dim3(kB), 0, b.st, ni, d_prefix, g, d_skey, d_sidx, d_pos, d_alpha, d_rad,
                     d_flag);
  RX_HIP(hipGetLastError());
  // the flagged particles in ascending order (the sort left d_idx = 0..n-1)
  RX_HIP(hipcub::DeviceSelect::Flagged(d_temp, t2, d_idx, d_flag, d_list, d_count, ni, b.st));
  hipLaunchKernelGGL(k_merge_resolve, dim3(1), dim3(kRT), 0, b.st, ni, d_list, d_count, g, d_skey, d_sidx, d_pos,
                     d_alpha, d_rad, d_vol, d_merges);
  RX_HIP(hipGetLastError());
  int64_t merges = 0;
  RX_HIP(down(&merges, d_merges, 1, b.st));
  RX_HIP(hipStreamSynchronize(b.st));
  if (merges > 0) {
    RX_HIP(down(pos, d_pos, 3 * nn, b.st));
    RX_HIP(down(alpha, d_alpha, 3 * nn, b.st));
    RX_HIP(down(volume, d_vol, nn, b.st));
    RX_HIP(hipStreamSynchronize(b.st));
  }
  if (nmerged) *nmerged = merges;
  return PW_OK;
}

}  // extern "C"
