// pw_internal.h -- device data layout and kernel launchers shared by the
// HIP translation units of libpawan_gpu.so.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>

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
// structure with {first, last, wake, -}; for their pair sums, nself = 1
// marks an item of a cross-wake block, which adds no self terms.)
struct WorkItem {
  int32_t t0, t1, s0, s1;
  int32_t part, nself;
  int32_t chunk0;  // source chunk (Chunk32) holding s0, for the scalar-fed FP32C kernel
  int32_t pad2;
};

// Sources as the scalar-fed FP32C kernel reads them, rebuilt from the state
// before each of its launches.  The sources of each system are cut into
// chunks of up to 64 consecutive particles; a chunk carries its bounding
// sphere, a source its position relative to its chunk's centre.  Both are
// read by whole waves at once (scalar loads), so they need no LDS staging.
struct alignas(16) Chunk32 {
  double cx, cy, cz;      // centre of the chunk's bounding box
  float r;                // radius of its bounding sphere (rounded up)
  float h2;               // largest sigma^2 / 2 of its sources
  int32_t first, count;   // its particles (set once, when the context is built)
};
struct alignas(16) Src32 {
  float rx, ry, rz;  // position relative to the chunk's centre (rounded once)
  float ax, ay, az;  // strength
  float mx, my, mz;  // a x r, in FP32 from the rounded a and r
  float hs2, vol, pad;
};
// Source record of the scalar-fed FP64 kernel: position relative to the
// chunk's centre, strength and m = a x r, in FP64 (80 bytes with padding).
struct alignas(16) Src64 {
  double rx, ry, rz, ax, ay, az, mx, my, mz, pad;
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
  Chunk32* chunks;    // scalar-fed FP32C: source chunks (geometry rebuilt per launch)
  Src32* src;         //   and source records, one per particle
  Src64* src64;       // scalar-fed FP64: source records relative to their chunk's centre
  int nchunks;
};

// How a right-hand side is launched (chosen by the autotuner).
struct PairConfig {
  int parts = 1;           // source ranges per target block
  bool persistent = false; // resident blocks pulling items from a queue
  bool waves8 = false;     // FP32C: register budget of 8 waves/SIMD
  bool scalar_feed = false;// FP32C: each wave streams the sources through
                           // scalar loads (Chunk32/Src32) instead of LDS
                           // tiles shared by the block's four waves
  bool two_per_lane = false;// FP32C scalar-fed: two targets per lane
};

// Pair-kernel variants.
enum PairKind : int {
  kPairFP64Exact = 0,  // the reference's kernel formula for every pair
  kPairFP64 = 1,       // same, closed far-field form where exp/erf terms
                       // are below FP64 rounding (rho^2 > 90 sigma^2)
  kPairFP32C = 2,      // FP32 pair math, compensated FP32 accumulation
  kPairFP64HO = 4,     // FP64, high-order algebraic kernel (highorder.h);
                       // launched in place of kinds 0 and 1 when a context
                       // uses that kernel (3 is the tree path's kind)
  kPairFP32CHO = 5,    // FP32C with that kernel, in place of kind 2
  kPairFP32CPlain = 6, // FP32 pairs summed in plain FP32, no compensation
                       // (pw_set_compensation(ctx, 0); one item per
                       // 256-target block, grid launch)
};

// Pair kernel of `kind` over args.items; for parts > 1 also the reduction
// of the partial sums into args.rates.  grid = resident blocks for the
// persistent form.
hipError_t launch_pair(int kind, const PairConfig& cfg, const PairArgs& args, int grid,
                       hipStream_t st);
// resident blocks per CU of the persistent pair kernel of `kind` (256-thread
// blocks: its waves per SIMD; the scalar-fed FP32C kernel's one-wave blocks:
// its waves per CU), from the occupancy calculator
int pair_kernel_waves(int kind, const PairConfig& cfg);
// resident one-wave blocks per CU of the tree path's near-field kernel
int near_kernel_blocks_per_cu(bool highorder, bool uniform_sigma);

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
// p[0, n) = 0 (a kernel, for the same reason)
hipError_t launch_zero_doubles(double* p, int64_t n, hipStream_t st);
// acc[i] += add[i] for i in [first, last)
hipError_t launch_add_rates(Rate64* acc, const Rate64* add, int first, int last, hipStream_t st);

// Smoothed vorticity field (__interaction::influence, interaction.cpp:
// 212-260): field[3 t + k] += sum over the item's sources s of
// zeta_sigma(|x_t - x_s|) a_s, sigma^2 = (s_s^2 + s_t^2)/2, with
// zeta_{s_t}(0) a_t for s = t (SELFINFLUENCE).  Items of one launch must
// have distinct targets.
hipError_t launch_vorticity_field(const Rec64* y, const WorkItem* items, int nitems, bool high_order,
                                  double* field, hipStream_t st);

// Diagnostics (FP64).  Linear sums: items[b] = {first, last, -, -} particle
// ranges inside one wake; out[11 * b + m] gets the block's sums of
// [a (3), x x a (3), x x (x x a) (3), |a| z (x^2+y^2), |a| (x^2+y^2)].
hipError_t launch_diag_linear(const Rec64* y, const WorkItem* items,
                              int nitems, double* out, hipStream_t st);
// Quadratic sums over the pair work items: out[kDiagQuad * b + m] gets the
// block's total of [helicity, KE, KEF, enstrophy, div-free enstrophy] rows
// (self terms included; the enstrophies are zero for the Gaussian kernel).
constexpr int kDiagQuad = 5;
hipError_t launch_diag_pairs(const Rec64* y, const double* radius,
                             const WorkItem* items, int nitems, double* out,
                             bool high_order, hipStream_t st);

// ---------------------------------------------------------------------
// Cluster tree of the learned far field.
// ---------------------------------------------------------------------
constexpr int kLeafSize = 256;  // particles per leaf at most (= one LDS tile)
// Leaves are the cells of one octree level (Morton-code prefix), cut into
// runs of at most 256: the level is the first whose mean occupancy is at
// most kCellTarget particles per non-empty cell.
constexpr int kCellTarget = 512;
constexpr int kFan = 8;         // children per node
constexpr int kMaxLevels = 12;  // 256 * 8^11 particles

// One node: particle range [first, last) of the Morton-sorted state, centre
// (particle mean), bounding radius r about it, and the vorticity moments
//   M = sum a, D[3i+j] = sum a_i delta_j, s0 = sum |a|,
//   S1 = sum |a| delta, s2 = sum |a| |delta|^2, sg = mean sigma
// with delta = x - centre; s1 and s3 bound sum |a| |delta| and
// sum |a| |delta|^3 (exact on leaves, shifted with the triangle inequality
// above).  child = index of the first child, nchild their count (level > 0).
struct alignas(16) NodeRec {
  double cx, cy, cz;
  float r, s0, s2, sg;
  float M[3];
  float D[9];
  float S1[3];
  int first, last, level, child;
  int nchild;
  float s1, s3;
};
static_assert(sizeof(NodeRec) == 128, "NodeRec is 128 B");

// Quadrupole moment of a node, Q^(p)_jl = sum a_p delta_j delta_l, 6 per
// component p in the order (jl) = xx, yy, zz, xy, xz, yz.
constexpr int kQuad = 18;
struct alignas(16) NodeQuad {
  float Q[kQuad];
  float pad[2];
};
static_assert(sizeof(NodeQuad) == 80, "NodeQuad is 80 B");

// Level layout of the node array: level L holds level_count[L] nodes from
// level_off[L]; level 0 are the leaves, in Morton order.
struct TreeShape {
  int n, nleaf, nnodes, nlevel, root;
  int level_off[kMaxLevels];
  int level_count[kMaxLevels];
};
TreeShape tree_shape(int n, int nleaf);

// Near-field work item: targets = leaf T, sources = T's near list [k0, k1);
// slot < 0: writes the rates, else partial-sum row `slot`.
struct NearItem {
  int T, k0, k1, slot;
};
constexpr int kHistNeed = 13;     // hist slots of the longest near and far lists
constexpr size_t kListLdsMax = 128 * 1024;  // LDS of the list kernel's level buffers at most
constexpr int kNearChunk = 128;   // near leaves per work item
constexpr int kNearMaxParts = 16; // work items per leaf at most
constexpr int kLeafChunks = kLeafSize / 64;  // 64-particle runs of a leaf

// Bounding sphere of a leaf's 64-particle run (runs from the leaf's first
// particle; r < 0 for a run past the leaf's end): the near kernel tests its
// 64-target slices against the sources' runs with these.
struct alignas(16) LeafChunk {
  double cx, cy, cz;
  float r;
  float pad;
};

struct TreeBuffers {
  double* bbox_part = nullptr;  // 6 x 1024 block extremes
  double* bbox = nullptr;       // lo[3], hi[3]
  uint32_t* code = nullptr;     // n Morton codes (original order)
  uint32_t* code_sorted = nullptr;
  int* idx = nullptr;           // 0..n-1
  int* perm = nullptr;          // sorted position -> original index
  void* sort_temp = nullptr;
  size_t sort_temp_bytes = 0;
  Rec64* sorted = nullptr;      // state in Morton order
  // leaves
  int* cell_start = nullptr;    // n: first sorted index of the particle's cell
  int* cut = nullptr;           // n: 1 where a leaf starts (scratch before)
  int* leaf_id = nullptr;       // n: exclusive scan of cut
  int* leaf_first = nullptr;    // leaf_cap + 1: first particle of each leaf, then n
  int* hist = nullptr;          // 16: Morton prefix-length histogram, [11] level, [12] nleaf,
                                //     [kHistNeed], [kHistNeed + 1]: longest near and far lists
  void* scan_n_temp = nullptr;
  size_t scan_n_bytes = 0;
  int leaf_cap = 0;             // leaves the leaf-indexed buffers hold
  int node_cap = 0;
  NodeRec* nodes = nullptr;
  NodeQuad* quad = nullptr;     // quadrupole of each node
  int* near_list = nullptr;     // nleaf x cap_near leaf ids
  int* near_cnt = nullptr;
  int cap_near = 0;
  int* far_list = nullptr;      // nleaf x cap_far node ids
  int* far_cnt = nullptr;
  int cap_far = 0;
  int* overflow = nullptr;      // set when a list or the stack overflows
  // near-field work items
  int* near_parts = nullptr;    // items per leaf
  int* near_offs = nullptr;     // first item of each leaf
  NearItem* near_items = nullptr;  // nleaf x kNearMaxParts at most
  int* near_total = nullptr;
  int* near_queue = nullptr;
  Rate64* near_partial = nullptr;  // item x 256 partial rates
  LeafChunk* leaf_chunks = nullptr;  // leaf x kLeafChunks run spheres
  void* scan_temp = nullptr;
  size_t scan_temp_bytes = 0;
};

// Traversal parameters: a node is far from target leaf T when
//   r_node + r_T < theta d,  (d - r_node - r_T)^2 > 44 sigma_max^2
//   and err_coef * s3 <= tol * (d - r_T)^5 (quad: the analytic expansion to
//   quadrupole order) or err_coef * s2 <= tol * (d - r_T)^4 (monopole +
//   dipole and the network's residual)
// (the last is the self-estimated truncation error of the far
// evaluation, err_coef calibrated for the evaluation in use).
struct ListParams {
  float theta;
  float tol;
  float err_coef;
  float sigma_max2;
  int quad;
  // high-order algebraic kernel: its velocity differs from the singular law
  // the far field expands by at most (15/8)(sigma/r)^4 of it, so a far node
  // also needs ho_coef s0 / gap^6 (ho_coef = (15/8) sigma_max^4 / 4 pi) added
  // to its self-estimated error to stay within tol, or, without a tolerance,
  // gap^2 > ho_gap2 sigma_max^2 (the deviation below FP32 rounding); both 0
  // for the Gaussian
  float ho_coef;
  float ho_gap2;
};

hipError_t tree_sort_temp_bytes(int n, size_t* bytes);
hipError_t tree_scan_temp_bytes(int n, size_t* bytes);
// bbox, Morton codes, radix sort, gather into tb.sorted; the leaves: octree
// level, cell starts (max-scan), cuts, leaf ids (exclusive scan),
// leaf_first (written only below leaf_cap) and the leaf count in hist[12].
hipError_t build_leaves(const Rec64* y, int n, TreeBuffers& tb, hipStream_t st);
// leaf_first again (after the leaf buffers have grown)
hipError_t write_leaf_first(int n, TreeBuffers& tb, hipStream_t st);
// leaf and parent moments
hipError_t build_moments(int n, TreeBuffers& tb, const TreeShape& shape, hipStream_t st);
// near and far lists of every target leaf
hipError_t build_lists(const TreeBuffers& tb, const TreeShape& shape, const ListParams& lp,
                       hipStream_t st);

// per target leaf: out[2T] = near pairs, out[2T+1] = far evaluations
hipError_t launch_list_stats(const TreeBuffers& tb, const TreeShape& shape, float sigma_max2, long long* out,
                             hipStream_t st);

// Near field over the lists: rates_s[t] = sum over the near leaves of
// target t's leaf (FP32C pair math, PSE included), Morton order.
// Near-field work items from the near lists (device only: parts per leaf,
// exclusive scan, items, total).
hipError_t near_scan_temp_bytes(int nleaf, size_t* bytes);
hipError_t build_near_items(const TreeBuffers& tb, int nleaf, hipStream_t st);
// Near field over the work items (persistent, `grid` resident blocks),
// then the reduction of split leaves.  Leaf pairs whose bounding spheres
// are more than sqrt(44) sigma_max apart skip the per-pair near-field test
// (no pair of them can need it).
// Raise a kernel's dynamic-LDS limit to `bytes` on the current device, once
// per device (hipFuncSetAttribute applies to the current device); `done`
// holds one bit per device id (ids >= 64 set the attribute on every call).
inline hipError_t ensure_lds_limit(const void* kernel, size_t bytes, std::atomic<unsigned long long>& done) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const unsigned long long bit = dev < 64 ? (1ull << dev) : 0ull;
  if (bit && (done.load(std::memory_order_acquire) & bit)) return hipSuccess;
  e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
  if (e == hipSuccess && bit) done.fetch_or(bit, std::memory_order_release);
  return e;
}

// Vortex-ring strength resolution (pw_vring.hip): influence matrix M (m x m)
// of the ring at `off` (nazi stations of m cells), its LU solve against
// `ideal` on a copy `lu`, and the strengths s written into the ring's
// vorticity.  flag holds 1 + m ints: flag[0] = 1 if the matrix is
// singular, the rest the row permutation.  m <= kVringMaxCells.
constexpr int kVringMaxCells = 1024;
hipError_t launch_vring_resolve(Rec64* y, int off, int m, int nazi, bool high_order, double* M,
                                double* lu, const double* ideal, double* s, int* flag, hipStream_t st);

// velocity at m points (AoS doubles, device) from the n particles of state
hipError_t launch_point_velocity(const Rec64* state, int n, const double* pts, int m, double* u,
                                 bool high_order, hipStream_t st);
// near field of the tree path (highorder: the algebraic kernel's pairs)
hipError_t launch_near_fp32c(const Rec64* sorted, const NodeRec* nodes, const TreeBuffers& tb, int nleaf,
                             double nu, float sigma_max2, bool uniform_sigma, bool highorder, int grid,
                             Rate64* rates_s, hipStream_t st);

// Far field over the lists added to rates_s: the analytic expansion to
// quadrupole order (quad != nullptr), or monopole + dipole and, with
// mlp != nullptr, the learned residual.
struct MlpWeights;  // packed BF16 fragments, pw_mlp.hip
hipError_t launch_far(const Rec64* sorted, const NodeRec* nodes, const NodeQuad* quad, const int* far_list,
                      const int* far_cnt, int cap_far, int nleaf, const MlpWeights* mlp,
                      Rate64* rates_s, hipStream_t st);
// the two passes of launch_far: the analytic expansion, and the network's
// residual added to it
hipError_t launch_far_md(const Rec64* sorted, const NodeRec* nodes, const NodeQuad* quad, const int* far_list,
                         const int* far_cnt, int cap_far, int nleaf, Rate64* rates_s, hipStream_t st);
hipError_t launch_far_mlp(const Rec64* sorted, const NodeRec* nodes, const int* far_list, const int* far_cnt,
                          int cap_far, int nleaf, const MlpWeights& mlp, Rate64* rates_s, hipStream_t st);

// dalpha_i -= mean over i of dalpha (part: 3 * (ceil(n/256) + 1) doubles)
hipError_t launch_remove_mean_dalpha(Rate64* rates, int n, double* part, hipStream_t st);
// rates[perm[i]] = rates_s[i]
hipError_t launch_scatter_rates(const Rate64* rates_s, const int* perm, int n, Rate64* rates,
                                hipStream_t st);

}  // namespace pw
