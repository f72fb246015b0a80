// pw_tree.hip -- the cluster tree of the learned far field.
//
// Particles are ordered along a Morton (Z-order) curve of their positions.
// Leaves are the non-empty cells of one octree level (the first level whose
// mean occupancy is at most kCellTarget), each cut into runs of at most 256
// consecutive particles: a leaf never straddles a jump of the curve, so its
// radius is bounded by its cell's.  (Fixed runs of 256 sorted particles do
// straddle such jumps; at N = 1M some then spanned the whole wake and were
// near to every other leaf.)  8 consecutive nodes of one level form a node
// of the next.  Every node stores its centre (particle mean), a bounding radius,
// and the vorticity moments the far field needs:
//   M  = sum a,   D_ij = sum a_i delta_j,   Q^(p)_jl = sum a_p delta_j delta_l
//   (delta = x - centre), s0 = sum |a|, S1 = sum |a| delta,
//   s_n = sum |a| |delta|^n (n = 1, 2, 3)
// (D, Q, S1, s2 of a parent from its children by the parallel-axis shift;
// s1 and s3 of a parent are bounds: |delta + h| <= |delta| + |h|).
// For each target leaf the tree is walked once per right-hand side:
// nodes that pass the opening-angle test (r_node + r_leaf < theta d), lie
// beyond the Gaussian core, and whose estimated far-evaluation error
// c s3 / (d - r_leaf)^5 for the analytic expansion to quadrupole order, or
// c s2 / (d - r_leaf)^4 for monopole + dipole and the network's residual
// (the first order the expansion leaves out; c is calibrated for the
// evaluation in use) is within the velocity tolerance
// go to the far list; leaves that fail go to the near list (evaluated by
// the direct sum), other nodes that fail are opened.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>

#include "pw_internal.h"

namespace pw {
namespace {

constexpr int kLeaf = kLeafSize;
// Far nodes keep every pair beyond rho^2 > 44 sigma^2, where the Gaussian
// terms of the kernel are below FP32 rounding.
constexpr float kFar32List = 44.0f;

// ---- bounding box -------------------------------------------------------
__global__ __launch_bounds__(256) void k_bbox_partial(const Rec64* __restrict__ y, int n,
                                                      double* __restrict__ part) {
  __shared__ double red[6][256];
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const Rec64 r = y[i];
    lo[0] = fmin(lo[0], r.x); lo[1] = fmin(lo[1], r.y); lo[2] = fmin(lo[2], r.z);
    hi[0] = fmax(hi[0], r.x); hi[1] = fmax(hi[1], r.y); hi[2] = fmax(hi[2], r.z);
  }
  for (int k = 0; k < 3; ++k) {
    red[k][threadIdx.x] = lo[k];
    red[3 + k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s)
      for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + s]);
        red[3 + k][threadIdx.x] = fmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; ++k) part[6 * blockIdx.x + k] = red[k][0];
}

__global__ __launch_bounds__(256) void k_bbox_final(const double* __restrict__ part, int nb, double* __restrict__ box) {
  __shared__ double red[6][4];
  double b[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int i = threadIdx.x; i < nb; i += 256)
    for (int k = 0; k < 3; ++k) {
      b[k] = fmin(b[k], part[6 * i + k]);
      b[3 + k] = fmax(b[3 + k], part[6 * i + 3 + k]);
    }
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < 3; ++k) {
      b[k] = fmin(b[k], __shfl_xor(b[k], off, 64));
      b[3 + k] = fmax(b[3 + k], __shfl_xor(b[3 + k], off, 64));
    }
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x >> 6] = b[k];
  __syncthreads();
  if (threadIdx.x < 3) {
    const int k = threadIdx.x;
    box[k] = fmin(fmin(red[k][0], red[k][1]), fmin(red[k][2], red[k][3]));
    box[3 + k] = fmax(fmax(red[3 + k][0], red[3 + k][1]), fmax(red[3 + k][2], red[3 + k][3]));
  }
}

// ---- Morton codes -------------------------------------------------------
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x30000ffu;
  v = (v | (v << 8)) & 0x300f00fu;
  v = (v | (v << 4)) & 0x30c30c3u;
  v = (v | (v << 2)) & 0x9249249u;
  return v;
}

__global__ __launch_bounds__(256) void k_morton(const Rec64* __restrict__ y, int n,
                                                const double* __restrict__ box,
                                                uint32_t* __restrict__ code, int* __restrict__ idx) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double span = fmax(fmax(box[3] - box[0], box[4] - box[1]), fmax(box[5] - box[2], 1e-300));
  const double s = 1023.0 / (span * 1.0000001);
  const Rec64 r = y[i];
  const uint32_t qx = static_cast<uint32_t>((r.x - box[0]) * s);
  const uint32_t qy = static_cast<uint32_t>((r.y - box[1]) * s);
  const uint32_t qz = static_cast<uint32_t>((r.z - box[2]) * s);
  code[i] = spread10(qx) | (spread10(qy) << 1) | (spread10(qz) << 2);
  idx[i] = i;
}

// Length, in 3-bit groups, of the common Morton prefix of sorted neighbours
// i-1 and i (10: same finest cell).
__device__ __forceinline__ int prefix_level(const uint32_t* __restrict__ code, int i) {
  const uint32_t x = code[i - 1] ^ code[i];
  if (x == 0) return 10;
  return (__clz(x) - 2) / 3;  // 30-bit codes: 2 leading zero bits always
}

__global__ __launch_bounds__(256) void k_prefix_hist(const uint32_t* __restrict__ code, int n, int* __restrict__ hist) {
  __shared__ int h[11];
  if (threadIdx.x < 11) h[threadIdx.x] = 0;
  __syncthreads();
  for (int i = 1 + blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) atomicAdd(&h[prefix_level(code, i)], 1);
  __syncthreads();
  if (threadIdx.x < 11 && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// the octree level of the leaves -> hist[11]
__global__ void k_leaf_level(int* __restrict__ hist, int n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  long long cells = 1;
  int level = 10;
  for (int D = 1; D <= 10; ++D) {
    cells += hist[D - 1];  // boundaries whose common prefix is shorter than D
    if (static_cast<long long>(n) <= static_cast<long long>(kCellTarget) * cells) {
      level = D;
      break;
    }
  }
  hist[11] = level;
}

// cell_start input: i where a cell of the leaf level starts, else 0
__global__ __launch_bounds__(256) void k_cell_marks(const uint32_t* __restrict__ code, int n,
                                                    const int* __restrict__ hist, int* __restrict__ mark) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int level = hist[11];
  mark[i] = (i == 0 || prefix_level(code, i) < level) ? i : 0;
}

// a leaf starts at every cell start and every 256 particles into a cell
__global__ __launch_bounds__(256) void k_cuts(const int* __restrict__ cell_start, int n, int* __restrict__ cut) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  cut[i] = ((i - cell_start[i]) % kLeaf) == 0 ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_leaf_first(const int* __restrict__ cut, const int* __restrict__ leaf_id, int n,
                                                    int leaf_cap, int* __restrict__ leaf_first, int* __restrict__ hist) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int id = leaf_id[i];
  if (cut[i] && id < leaf_cap) leaf_first[id] = i;
  if (i == n - 1) {
    const int nleaf = id + cut[i];
    hist[12] = nleaf;
    if (nleaf <= leaf_cap) leaf_first[nleaf] = n;
  }
}

__global__ __launch_bounds__(256) void k_gather(const Rec64* __restrict__ y, const int* __restrict__ perm,
                                                int n, Rec64* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = y[perm[i]];
}

// ---- leaf moments: one block per leaf -----------------------------------
// Deterministic block sum of M doubles (wave butterfly, then the 4 wave
// totals in a fixed order).
template <int M>
__device__ __forceinline__ void block_sum_d(double (&v)[M], double* red) {
  for (int off = 32; off > 0; off >>= 1)
    for (int m = 0; m < M; ++m) v[m] += __shfl_xor(v[m], off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
    for (int m = 0; m < M; ++m) red[wave * M + m] = v[m];
  __syncthreads();
  for (int m = 0; m < M; ++m) v[m] = (red[m] + red[M + m]) + (red[2 * M + m] + red[3 * M + m]);
  __syncthreads();
}

__global__ __launch_bounds__(kLeaf) void k_leaf_moments(const Rec64* __restrict__ ys,
                                                        const int* __restrict__ leaf_first,
                                                        NodeRec* __restrict__ nodes, NodeQuad* __restrict__ quad) {
  __shared__ double red[4 * 20];
  const int leaf = blockIdx.x;
  const int first = leaf_first[leaf], last = leaf_first[leaf + 1];
  const int i = first + static_cast<int>(threadIdx.x);
  const bool on = i < last;
  const Rec64 p = ys[on ? i : first];
  double c[3] = {on ? p.x : 0.0, on ? p.y : 0.0, on ? p.z : 0.0};
  block_sum_d<3>(c, red);
  const double cnt = static_cast<double>(last - first);
  const double cx = c[0] / cnt, cy = c[1] / cnt, cz = c[2] / cnt;
  const double dx = p.x - cx, dy = p.y - cy, dz = p.z - cz;
  const double am = sqrt(p.ax * p.ax + p.ay * p.ay + p.az * p.az);
  double v[18];
  if (on) {
    v[0] = p.ax; v[1] = p.ay; v[2] = p.az;
    v[3] = p.ax * dx; v[4] = p.ax * dy; v[5] = p.ax * dz;
    v[6] = p.ay * dx; v[7] = p.ay * dy; v[8] = p.ay * dz;
    v[9] = p.az * dx; v[10] = p.az * dy; v[11] = p.az * dz;
    v[12] = am; v[13] = am * dx; v[14] = am * dy; v[15] = am * dz;
    v[16] = am * (dx * dx + dy * dy + dz * dz);
    v[17] = sqrt(p.s2);
  } else {
    for (int m = 0; m < 18; ++m) v[m] = 0.0;
  }
  block_sum_d<18>(v, red);
  // quadrupole (a_p delta_j delta_l, (jl) = xx yy zz xy xz yz), s1, s3
  double q[20];
  if (on) {
    const double dd[6] = {dx * dx, dy * dy, dz * dz, dx * dy, dx * dz, dy * dz};
    const double ap[3] = {p.ax, p.ay, p.az};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int m = 0; m < 6; ++m) q[6 * a + m] = ap[a] * dd[m];
    const double dn = sqrt(dx * dx + dy * dy + dz * dz);
    q[18] = am * dn;
    q[19] = am * dn * dn * dn;
  } else {
    for (int m = 0; m < 20; ++m) q[m] = 0.0;
  }
  block_sum_d<20>(q, red);
  // bounding radius about the centre
  double d = on ? sqrt(dx * dx + dy * dy + dz * dz) : 0.0;
  for (int off = 32; off > 0; off >>= 1) d = fmax(d, __shfl_xor(d, off, 64));
  __shared__ double rmax[4];
  if ((threadIdx.x & 63) == 0) rmax[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    NodeRec nd;
    nd.cx = cx; nd.cy = cy; nd.cz = cz;
    // rounded up so that the radius stays a bound in FP32
    nd.r = static_cast<float>(fmax(fmax(rmax[0], rmax[1]), fmax(rmax[2], rmax[3])) * (1.0 + 1e-6));
    for (int m = 0; m < 3; ++m) nd.M[m] = static_cast<float>(v[m]);
    for (int m = 0; m < 9; ++m) nd.D[m] = static_cast<float>(v[3 + m]);
    nd.s0 = static_cast<float>(v[12]);
    for (int m = 0; m < 3; ++m) nd.S1[m] = static_cast<float>(v[13 + m]);
    nd.s2 = static_cast<float>(v[16]);
    nd.sg = static_cast<float>(v[17] / cnt);
    nd.first = first; nd.last = last;
    nd.level = 0; nd.child = -1;
    nd.nchild = 0;
    nd.s1 = static_cast<float>(q[18]);
    nd.s3 = static_cast<float>(q[19]);
    nodes[leaf] = nd;
    NodeQuad nq;
    for (int m = 0; m < kQuad; ++m) nq.Q[m] = static_cast<float>(q[m]);
    nq.pad[0] = nq.pad[1] = 0.0f;
    quad[leaf] = nq;
  }
}

// ---- parent moments: one thread per node of a level ----------------------
// Moments about the parent centre from the children's by the parallel-axis
// shift h = c_child - c_parent:
//   D += M_c h^T,  S1 += S1_c + s0_c h,  s2 += s2_c + 2 S1_c.h + s0_c |h|^2,
//   Q^(p)_jl += Q_c^(p)_jl + D_c[p][j] h_l + D_c[p][l] h_j + M_c[p] h_j h_l,
//   s1 += s1_c + s0_c |h|,  s3 += s3_c + 3 s2_c |h| + 3 s1_c |h|^2 + s0_c |h|^3
//   (bounds, then s1 <= r s0 and s3 <= r s2), r = max(|h| + r_c).
__global__ __launch_bounds__(64) void k_parent_moments(NodeRec* __restrict__ nodes, NodeQuad* __restrict__ quad,
                                                       int child_off, int nchild, int off, int count, int level) {
  const int j = blockIdx.x * 64 + threadIdx.x;
  if (j >= count) return;
  const int c0 = child_off + kFan * j, c1 = child_off + min(kFan * (j + 1), nchild);
  double sx = 0, sy = 0, sz = 0, cnt = 0, sg = 0;
  for (int c = c0; c < c1; ++c) {
    const double w = nodes[c].last - nodes[c].first;
    sx += w * nodes[c].cx; sy += w * nodes[c].cy; sz += w * nodes[c].cz; cnt += w;
    sg += w * nodes[c].sg;
  }
  NodeRec p;
  p.cx = sx / cnt; p.cy = sy / cnt; p.cz = sz / cnt;
  double M[3] = {0, 0, 0}, D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, s0 = 0, S1[3] = {0, 0, 0}, s2 = 0, r = 0;
  double Q[kQuad], s1 = 0, s3 = 0;
  for (int m = 0; m < kQuad; ++m) Q[m] = 0.0;
  constexpr int kJ[6] = {0, 1, 2, 0, 0, 1}, kL[6] = {0, 1, 2, 1, 2, 2};  // (jl) of each Q slot
  for (int c = c0; c < c1; ++c) {
    const NodeRec& q = nodes[c];
    const NodeQuad& qq = quad[c];
    const double h[3] = {q.cx - p.cx, q.cy - p.cy, q.cz - p.cz};
    double h2 = 0, s1h = 0;
    for (int a = 0; a < 3; ++a) {
      M[a] += q.M[a];
      for (int b = 0; b < 3; ++b) D[3 * a + b] += q.D[3 * a + b] + static_cast<double>(q.M[a]) * h[b];
      for (int m = 0; m < 6; ++m) {
        const int j = kJ[m], l = kL[m];
        Q[6 * a + m] += qq.Q[6 * a + m] + static_cast<double>(q.D[3 * a + j]) * h[l] +
                        static_cast<double>(q.D[3 * a + l]) * h[j] + static_cast<double>(q.M[a]) * h[j] * h[l];
      }
      S1[a] += q.S1[a] + static_cast<double>(q.s0) * h[a];
      h2 += h[a] * h[a];
      s1h += q.S1[a] * h[a];
    }
    const double hn = sqrt(h2);
    s0 += q.s0;
    s2 += q.s2 + 2.0 * s1h + q.s0 * h2;
    s1 += q.s1 + q.s0 * hn;
    s3 += q.s3 + 3.0 * q.s2 * hn + 3.0 * q.s1 * h2 + q.s0 * h2 * hn;
    r = fmax(r, hn + q.r);
  }
  p.r = static_cast<float>(r * (1.0 + 1e-6));
  p.s1 = static_cast<float>(fmin(s1, r * s0));
  p.s3 = static_cast<float>(fmin(s3, r * fmax(s2, 0.0)));
  NodeQuad pq;
  for (int m = 0; m < kQuad; ++m) pq.Q[m] = static_cast<float>(Q[m]);
  pq.pad[0] = pq.pad[1] = 0.0f;
  quad[off + j] = pq;
  p.s0 = static_cast<float>(s0);
  p.s2 = static_cast<float>(s2);
  p.sg = static_cast<float>(sg / cnt);
  for (int m = 0; m < 3; ++m) p.M[m] = static_cast<float>(M[m]);
  for (int m = 0; m < 9; ++m) p.D[m] = static_cast<float>(D[m]);
  for (int m = 0; m < 3; ++m) p.S1[m] = static_cast<float>(S1[m]);
  p.first = nodes[c0].first;
  p.last = nodes[c1 - 1].last;
  p.level = level;
  p.child = c0;
  p.nchild = c1 - c0;
  nodes[off + j] = p;
}

// ---- interaction lists: one wave per target leaf --------------------------
// Breadth-first from the root, one level per pass: the wave tests 64
// candidates at a time (the children of the nodes opened on the level
// above, taken in order) and appends far nodes, near leaves and opened
// nodes in lane order through ballots, so the lists are deterministic: far
// nodes coarse level first, each level in Morton order, and the near leaves
// (all on the last level) in Morton order.  The nodes opened on one level
// wait in LDS, two buffers of max_open entries (the largest internal level,
// so they cannot overflow).  need[0], need[1] collect the longest near and
// far lists, which may exceed the capacities: the lists are then cut (and
// the overflow flag set), and the host grows them and builds again.
__device__ __forceinline__ bool is_far(const NodeRec& q, double tcx, double tcy, double tcz, float tr,
                                       const ListParams& lp) {
  const double ddx = q.cx - tcx, ddy = q.cy - tcy, ddz = q.cz - tcz;
  const float d = static_cast<float>(sqrt(ddx * ddx + ddy * ddy + ddz * ddz));
  const float gap = d - q.r - tr;
  bool far = (q.r + tr) < lp.theta * d && gap > 0.0f && gap * gap > kFar32List * lp.sigma_max2;
  if (far && lp.ho_gap2 > 0.0f) far = gap * gap > lp.ho_gap2 * lp.sigma_max2;
  if (far && lp.tol > 0.0f) {
    const float dm = d - tr;
    // the first order the far evaluation leaves out
    const float dmn = lp.quad ? (dm * dm) * (dm * dm) * dm : (dm * dm) * (dm * dm);
    const float sn = lp.quad ? q.s3 : q.s2;
    if (lp.ho_coef > 0.0f) {
      const float g2 = gap * gap;
      far = lp.err_coef * sn / dmn + lp.ho_coef * q.s0 / (g2 * g2 * g2) <= lp.tol;
    } else {
      far = lp.err_coef * sn <= lp.tol * dmn;
    }
  }
  return far;
}

__global__ __launch_bounds__(64) void k_lists(const NodeRec* __restrict__ nodes, TreeShape shape,
                                              ListParams lp, int* __restrict__ near_list,
                                              int* __restrict__ near_cnt, int cap_near,
                                              int* __restrict__ far_list, int* __restrict__ far_cnt,
                                              int cap_far, int max_open, int* __restrict__ overflow,
                                              int* __restrict__ need) {
  extern __shared__ int open_lds[];
  int* open[2] = {open_lds, open_lds + max_open};
  const int T = blockIdx.x;
  const int lane = static_cast<int>(threadIdx.x);
  const double tcx = nodes[T].cx, tcy = nodes[T].cy, tcz = nodes[T].cz;
  const float tr = nodes[T].r;
  int* __restrict__ NL = near_list + static_cast<size_t>(T) * cap_near;
  int* __restrict__ FL = far_list + static_cast<size_t>(T) * cap_far;
  const unsigned long long below = (1ull << lane) - 1ull;
  int nn = 0, nf = 0;
  bool over = false;
  // level 0 of the walk: the root alone, tested as the only candidate
  int np = 0, cur = 0;
  {
    const int id = shape.root;
    if (is_far(nodes[id], tcx, tcy, tcz, tr, lp)) {
      if (lane == 0) FL[0] = id;  // (cap_far >= 1)
      nf = 1;
    } else if (nodes[id].level == 0) {
      if (lane == 0) NL[0] = id;
      nn = 1;
    } else {
      if (lane == 0) open[0][0] = id;
      np = 1;
    }
  }
  while (np > 0) {
    __syncthreads();  // the opened nodes of the level above are in open[cur]
    int nnext = 0;
    for (int base = 0; base < kFan * np; base += 64) {
      const int s = base + lane, j = s / kFan, k = s % kFan;
      int id = -1;
      if (j < np) {
        const NodeRec& P = nodes[open[cur][j]];
        if (k < P.nchild) id = P.child + k;
      }
      bool far = false, near = false, opn = false;
      if (id >= 0) {
        const NodeRec& q = nodes[id];
        far = is_far(q, tcx, tcy, tcz, tr, lp);
        near = !far && q.level == 0;
        opn = !far && !near;
      }
      const unsigned long long mf = __ballot(far), mn = __ballot(near), mo = __ballot(opn);
      if (far) {
        const int p = nf + __popcll(mf & below);
        if (p < cap_far) FL[p] = id;
      }
      if (near) {
        const int p = nn + __popcll(mn & below);
        if (p < cap_near) NL[p] = id;
      }
      if (opn) {
        const int p = nnext + __popcll(mo & below);
        if (p < max_open) open[cur ^ 1][p] = id;
      }
      nf += __popcll(mf);
      nn += __popcll(mn);
      nnext += __popcll(mo);
    }
    if (nnext > max_open) {  // (cannot happen: a level has at most max_open nodes)
      over = true;
      nnext = max_open;
    }
    cur ^= 1;
    np = nnext;
  }
  if (nn > cap_near || nf > cap_far) over = true;
  if (lane == 0) {
    near_cnt[T] = min(nn, cap_near);
    far_cnt[T] = min(nf, cap_far);
    atomicMax(need, nn);
    atomicMax(need + 1, nf);
    if (over) atomicOr(overflow, 1);
  }
}

__global__ void k_fill_zero16(int* __restrict__ p) {
  if (threadIdx.x < 16) p[threadIdx.x] = 0;
}

// Near-field work items: parts per leaf, then the items.
__global__ __launch_bounds__(256) void k_near_parts(const int* __restrict__ near_cnt, int nleaf,
                                                    int* __restrict__ parts) {
  const int T = blockIdx.x * 256 + threadIdx.x;
  if (T >= nleaf) return;
  const int nc = near_cnt[T];
  parts[T] = min(kNearMaxParts, max(1, (nc + kNearChunk - 1) / kNearChunk));
}

__global__ __launch_bounds__(256) void k_near_emit(const int* __restrict__ near_cnt, const int* __restrict__ parts,
                                                   const int* __restrict__ offs, int nleaf,
                                                   NearItem* __restrict__ items, int* __restrict__ total) {
  const int T = blockIdx.x * 256 + threadIdx.x;
  if (T >= nleaf) return;
  const int nc = near_cnt[T], P = parts[T], o = offs[T];
  for (int p = 0; p < P; ++p) {
    NearItem it;
    it.T = T;
    it.k0 = static_cast<int>(static_cast<long long>(nc) * p / P);
    it.k1 = static_cast<int>(static_cast<long long>(nc) * (p + 1) / P);
    it.slot = P == 1 ? -1 : o + p;
    items[o + p] = it;
  }
  if (T == nleaf - 1) *total = o + P;
}

// Bounding sphere of every 64-particle run of every leaf, one wave per run:
// the centre of the run's bounding box and the largest distance from it
// (FP64, rounded up to FP32).
__device__ __forceinline__ double wave_reduce_min(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ double wave_reduce_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

__global__ __launch_bounds__(64) void k_leaf_chunks(const Rec64* __restrict__ sorted,
                                                    const NodeRec* __restrict__ nodes,
                                                    LeafChunk* __restrict__ out) {
  const int L = blockIdx.x / kLeafChunks, c = blockIdx.x % kLeafChunks;
  const int first = nodes[L].first + 64 * c, last = nodes[L].last;
  const int lane = static_cast<int>(threadIdx.x);
  if (first >= last) {  // the same for every lane
    if (lane == 0) out[blockIdx.x] = LeafChunk{0.0, 0.0, 0.0, -1.0f, 0.0f};
    return;
  }
  const int i = first + lane < last ? first + lane : first;
  const double x = sorted[i].x, y = sorted[i].y, z = sorted[i].z;
  const double cx = 0.5 * (wave_reduce_min(x) + wave_reduce_max(x));
  const double cy = 0.5 * (wave_reduce_min(y) + wave_reduce_max(y));
  const double cz = 0.5 * (wave_reduce_min(z) + wave_reduce_max(z));
  const double d2 = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz);
  const double r = sqrt(wave_reduce_max(d2));
  if (lane == 0) out[blockIdx.x] = LeafChunk{cx, cy, cz, static_cast<float>(r * (1.0 + 1e-6)), 0.0f};
}

// Per target leaf: near pairs (targets x near sources, self pairs
// excluded), far evaluations (targets x far nodes), and the near pairs of
// listed leaves within the Gaussian core range (bounding spheres at most
// sqrt(44) sigma_max apart: the near kernel tests these pair by pair).
__global__ __launch_bounds__(64) void k_list_stats(const NodeRec* __restrict__ nodes, int nleaf,
                                                   const int* __restrict__ near_list,
                                                   const int* __restrict__ near_cnt, int cap_near,
                                                   const int* __restrict__ far_cnt, float gap2,
                                                   long long* __restrict__ out) {
  const int T = blockIdx.x * 64 + threadIdx.x;
  if (T >= nleaf) return;
  const long long nt = nodes[T].last - nodes[T].first;
  long long src = 0, core = 0;
  for (int k = 0; k < near_cnt[T]; ++k) {
    const int id = near_list[static_cast<size_t>(T) * cap_near + k];
    const long long ns = nodes[id].last - nodes[id].first;
    src += ns;
    const double ddx = nodes[id].cx - nodes[T].cx, ddy = nodes[id].cy - nodes[T].cy, ddz = nodes[id].cz - nodes[T].cz;
    const float gap = static_cast<float>(sqrt(ddx * ddx + ddy * ddy + ddz * ddz)) - nodes[T].r - nodes[id].r;
    if (!(gap > 0.0f && gap * gap > gap2 * 1.0001f)) core += ns;
  }
  out[3 * T] = nt * src - nt;  // the leaf itself is on its near list
  out[3 * T + 1] = nt * far_cnt[T];
  out[3 * T + 2] = nt * core - nt;
}

}  // namespace

hipError_t launch_list_stats(const TreeBuffers& tb, const TreeShape& shape, float sigma_max2, long long* out,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_list_stats, dim3((shape.nleaf + 63) / 64), dim3(64), 0, st, tb.nodes, shape.nleaf,
                     tb.near_list, tb.near_cnt, tb.cap_near, tb.far_cnt, kFar32List * sigma_max2, out);
  return hipGetLastError();
}

hipError_t near_scan_temp_bytes(int nleaf, size_t* bytes) {
  return hipcub::DeviceScan::ExclusiveSum(nullptr, *bytes, static_cast<int*>(nullptr), static_cast<int*>(nullptr),
                                          nleaf);
}

hipError_t build_near_items(const TreeBuffers& tb, int nleaf, hipStream_t st) {
  const int nb = (nleaf + 255) / 256;
  hipLaunchKernelGGL(k_near_parts, dim3(nb), dim3(256), 0, st, tb.near_cnt, nleaf, tb.near_parts);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t bytes = tb.scan_temp_bytes;
  e = hipcub::DeviceScan::ExclusiveSum(tb.scan_temp, bytes, tb.near_parts, tb.near_offs, nleaf, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_near_emit, dim3(nb), dim3(256), 0, st, tb.near_cnt, tb.near_parts, tb.near_offs, nleaf,
                     tb.near_items, tb.near_total);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_leaf_chunks, dim3(nleaf * kLeafChunks), dim3(64), 0, st, tb.sorted, tb.nodes,
                     tb.leaf_chunks);
  return hipGetLastError();
}

TreeShape tree_shape(int n, int nleaf) {
  TreeShape s{};
  s.n = n;
  s.nleaf = nleaf;
  int count = s.nleaf, off = 0, L = 0;
  for (;;) {
    s.level_off[L] = off;
    s.level_count[L] = count;
    off += count;
    ++L;
    if (count == 1 || L == kMaxLevels) break;
    count = (count + kFan - 1) / kFan;
  }
  s.nlevel = L;
  s.nnodes = off;
  s.root = s.level_off[L - 1];
  return s;
}

hipError_t tree_sort_temp_bytes(int n, size_t* bytes) {
  return hipcub::DeviceRadixSort::SortPairs(nullptr, *bytes, static_cast<uint32_t*>(nullptr),
                                            static_cast<uint32_t*>(nullptr), static_cast<int*>(nullptr),
                                            static_cast<int*>(nullptr), n, 0, 30);
}

hipError_t tree_scan_temp_bytes(int n, size_t* bytes) {
  size_t a = 0, b = 0;
  hipError_t e = hipcub::DeviceScan::InclusiveScan(nullptr, a, static_cast<int*>(nullptr), static_cast<int*>(nullptr),
                                                   hipcub::Max(), n);
  if (e != hipSuccess) return e;
  e = hipcub::DeviceScan::ExclusiveSum(nullptr, b, static_cast<int*>(nullptr), static_cast<int*>(nullptr), n);
  *bytes = a > b ? a : b;
  return e;
}

hipError_t build_leaves(const Rec64* y, int n, TreeBuffers& tb, hipStream_t st) {
  const int nb = min(1024, (n + 255) / 256);
  const int grid = (n + 255) / 256;
  hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(256), 0, st, y, n, tb.bbox_part);
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(256), 0, st, tb.bbox_part, nb, tb.bbox);
  hipLaunchKernelGGL(k_morton, dim3(grid), dim3(256), 0, st, y, n, tb.bbox, tb.code, tb.idx);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t bytes = tb.sort_temp_bytes;
  e = hipcub::DeviceRadixSort::SortPairs(tb.sort_temp, bytes, tb.code, tb.code_sorted, tb.idx,
                                         tb.perm, n, 0, 30, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_gather, dim3(grid), dim3(256), 0, st, y, tb.perm, n, tb.sorted);
  hipLaunchKernelGGL(k_fill_zero16, dim3(1), dim3(64), 0, st, tb.hist);
  hipLaunchKernelGGL(k_prefix_hist, dim3(nb), dim3(256), 0, st, tb.code_sorted, n, tb.hist);
  hipLaunchKernelGGL(k_leaf_level, dim3(1), dim3(64), 0, st, tb.hist, n);
  hipLaunchKernelGGL(k_cell_marks, dim3(grid), dim3(256), 0, st, tb.code_sorted, n, tb.hist, tb.cut);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  bytes = tb.scan_n_bytes;
  e = hipcub::DeviceScan::InclusiveScan(tb.scan_n_temp, bytes, tb.cut, tb.cell_start, hipcub::Max(), n, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_cuts, dim3(grid), dim3(256), 0, st, tb.cell_start, n, tb.cut);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  bytes = tb.scan_n_bytes;
  e = hipcub::DeviceScan::ExclusiveSum(tb.scan_n_temp, bytes, tb.cut, tb.leaf_id, n, st);
  if (e != hipSuccess) return e;
  return write_leaf_first(n, tb, st);
}

hipError_t write_leaf_first(int n, TreeBuffers& tb, hipStream_t st) {
  hipLaunchKernelGGL(k_leaf_first, dim3((n + 255) / 256), dim3(256), 0, st, tb.cut, tb.leaf_id, n, tb.leaf_cap,
                     tb.leaf_first, tb.hist);
  return hipGetLastError();
}

hipError_t build_moments(int n, TreeBuffers& tb, const TreeShape& shape, hipStream_t st) {
  (void)n;
  hipLaunchKernelGGL(k_leaf_moments, dim3(shape.nleaf), dim3(kLeaf), 0, st, tb.sorted, tb.leaf_first, tb.nodes,
                     tb.quad);
  for (int L = 1; L < shape.nlevel; ++L)
    hipLaunchKernelGGL(k_parent_moments, dim3((shape.level_count[L] + 63) / 64), dim3(64), 0, st,
                       tb.nodes, tb.quad, shape.level_off[L - 1], shape.level_count[L - 1], shape.level_off[L],
                       shape.level_count[L], L);
  return hipGetLastError();
}

hipError_t build_lists(const TreeBuffers& tb, const TreeShape& shape, const ListParams& lp,
                       hipStream_t st) {
  if (shape.nleaf <= 0) return hipSuccess;
  int max_open = 1;
  for (int L = 1; L < shape.nlevel; ++L) max_open = std::max(max_open, shape.level_count[L]);
  const size_t lds = 2 * static_cast<size_t>(max_open) * sizeof(int);
  if (lds > kListLdsMax) return hipErrorInvalidValue;
  if (lds > 64 * 1024) {  // beyond the default limit
    static std::atomic<unsigned long long> lds_raised{0};
    const hipError_t e = ensure_lds_limit(reinterpret_cast<const void*>(k_lists), kListLdsMax, lds_raised);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_lists, dim3(shape.nleaf), dim3(64), lds, st, tb.nodes, shape, lp, tb.near_list,
                     tb.near_cnt, tb.cap_near, tb.far_list, tb.far_cnt, tb.cap_far, max_open, tb.overflow,
                     tb.hist + kHistNeed);
  return hipGetLastError();
}

}  // namespace pw
