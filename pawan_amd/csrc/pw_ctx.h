// pw_ctx.h -- the solver context behind the opaque pw_ctx handle, and the
// error plumbing shared by the translation units that implement the
// extern "C" entry points.  Internal to libpawan_gpu.so.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/pawan_gpu.h"
#include "pw_internal.h"
#include "pw_mlp.h"

namespace pwi {

// records msg as the thread's last error and returns code
int fail(int code, const std::string& msg);

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

constexpr int kTargetsPerItem = 256;
// pair-kernel kinds 0..2 are pw::PairKind; the tree path is kind 3
constexpr int kKindLearned = 3;

// State of the learned far field (tree path) of one context.
struct FarField {
  bool ready = false;
  pw::TreeShape shape{};
  pw::TreeBuffers tb;
  pw::Rate64* rates_s = nullptr;  // rates in Morton order
  int* d_overflow = nullptr;
  float sigma_max2 = 0.0f;
  bool uniform_sigma = false;     // every particle of one radius (the near field's constant core factors)
  // traversal parameters
  float theta = 0.5f;
  float tol = 0.0f;                  // absolute velocity tolerance (<= 0: opening angle only)
  bool use_mlp = true;
  bool conserve = true;              // remove the mean vorticity rate
  double* d_dpart = nullptr;         // block sums for it
  double coef_md = 0.3183098862;     // analytic path: 4/(4 pi), the octupole's bound, until calibrated
  double coef_mlp = 0.2387324146;    // network path: 3/(4 pi), the quadrupole's bound, until calibrated
  // network
  bool have_weights = false;
  std::vector<float> h_params;
  float *d_params = nullptr, *d_m = nullptr, *d_v = nullptr, *d_grad = nullptr;
  float *d_gpart = nullptr, *d_lpart = nullptr, *d_loss = nullptr;
  int gpart_blocks = 0;
  uint16_t* d_frag = nullptr;
  float* d_bias = nullptr;
  int adam_step = 0;
  // training samples
  float *d_X = nullptr, *d_Y = nullptr;
  double* d_aux = nullptr;
  int64_t ns = 0, ns_cap = 0;
  // last lists built
  bool lists_valid = false;
  // stage timing of the right-hand sides while the context profiles: tree
  // and lists, near field, analytic far field, network, the rest
  static constexpr int kStages = 5;
  std::vector<std::array<hipEvent_t, kStages + 1>> st_used, st_free;
  double stage_ms[kStages] = {0, 0, 0, 0, 0};
  int64_t stage_rhs = 0;
  // network evaluations (far-list entries) of the last lists
  int64_t mlp_evals = 0;
};

}  // namespace pwi

struct pw_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  double nu = 1.56e-5;  // interaction.cpp:15
  bool coupled = true;
  int smoothing = PW_KERNEL_GAUSSIAN;  // PW_KERNEL_*
  bool compensated = true;             // FP32C sums (pw_set_compensation)

  // host description of the wakes
  std::vector<int64_t> wn, wmax, woff{0};
  std::vector<double> radius, volume;   // per particle
  std::vector<pw::Rec64> pending;       // state until the device is built
  bool built = false;
  int64_t n = 0;

  // device buffers
  pw::Rec64* y = nullptr;    // current state
  pw::Rec64* ys = nullptr;   // stage state
  pw::Rate64* k[4] = {nullptr, nullptr, nullptr, nullptr};
  pw::Rate64* rates = nullptr;  // rates of the last right-hand side
  double* d_radius = nullptr;
  pw::WorkItem* pair_items = nullptr;
  int npair = 0;
  pw::WorkItem* lin_items = nullptr;
  int nlin = 0;
  std::vector<pw::WorkItem> h_pair_items, h_lin_items;
  std::vector<int> pair_sys, lin_wake;  // system / wake of each item
  double* d_lin = nullptr;
  double* d_quad = nullptr;
  int* d_flag = nullptr;

  // per pair-kernel kind: launch configuration and its work items
  struct KindState {
    pw::PairConfig cfg;
    pw::WorkItem* d_items = nullptr;
    int nitems = 0;
    int built_parts = 0;  // parts the items were built for (0: none yet)
    bool tuned = false;
  } ks[3];
  pw::Rate64* partial = nullptr;  // source-split partial rates, parts x n
  int partial_parts = 0;
  int* d_queue = nullptr;     // work counter of the persistent kernels
  // source chunks of the scalar-fed FP32C kernel: runs of up to 64
  // particles of one system (chunk_first: their first particles, ascending)
  std::vector<int32_t> chunk_first;
  pw::Chunk32* d_chunks = nullptr;
  pw::Src32* d_src32 = nullptr;
  pw::Src64* d_src64 = nullptr;  // scalar-fed FP64 records (allocated on first use)
  int nchunks = 0;
  int num_cus = 256;
  struct SysRange {
    int64_t a, b;
    int id;
  };
  std::vector<SysRange> systems;  // largest first

  // HIP graphs of one step, per (scheme, kind, dt)
  bool use_graph = false;
  std::map<std::tuple<int, int, uint64_t>, hipGraphExec_t> graphs;

  std::vector<double> zc;  // persistent centroid per system

  // learned far field
  pwi::FarField ff;

  // profiling
  bool profiling = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_used, ev_free;
  double prof_ms = 0.0;
  int64_t prof_launches = 0;

  int nsys() const { return coupled ? 1 : static_cast<int>(wn.size()); }
};

namespace pwi {

int ensure_built(pw_ctx* c);
// event pair for one profiled launch (recycled)
int prof_begin(pw_ctx* c, hipEvent_t* e0, hipEvent_t* e1);
// learned far field: one right-hand side, list-overflow check, release
int ff_rhs(pw_ctx* c, const pw::Rec64* state, pw::Rate64* out);
int ff_check(pw_ctx* c);
void ff_release(pw_ctx* c);
// fold the recorded stage events into the totals (synchronises)
int ff_collect_stages(pw_ctx* c);
void ff_reset_stages(pw_ctx* c);

}  // namespace pwi

#define PW_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return ::pwi::fail(PW_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define PW_CTX(c)                                                   \
  do {                                                              \
    if (!(c)) return ::pwi::fail(PW_ERR_ARG, "null context");       \
    if (hipSetDevice((c)->device) != hipSuccess)                    \
      return ::pwi::fail(PW_ERR_HIP, "hipSetDevice failed");        \
  } while (0)

#define PW_TRY(expr)                     \
  do {                                   \
    int rc_ = (expr);                    \
    if (rc_ != PW_OK) return rc_;        \
  } while (0)
