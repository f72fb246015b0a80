// pw_ff.hip -- the learned far field (PW_LEARNED) of a context: tree and
// list buffers, the right-hand side over them, training-sample generation,
// network training and calibration, and the pw_ff_* entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "pw_ctx.h"

using pw::Rate64;
using pw::Rec64;
using pwi::dalloc;
using pwi::fail;

namespace {

constexpr double kQ = 0x1.45f306dc9c883p-4;  // 1/(4 pi)

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void free_leaf_buffers(pwi::FarField& f) {
  auto& tb = f.tb;
  for (void* p : {static_cast<void*>(tb.leaf_first), static_cast<void*>(tb.nodes), static_cast<void*>(tb.quad),
                  static_cast<void*>(tb.near_list),
                  static_cast<void*>(tb.near_cnt), static_cast<void*>(tb.far_list), static_cast<void*>(tb.far_cnt),
                  static_cast<void*>(tb.near_parts), static_cast<void*>(tb.near_offs),
                  static_cast<void*>(tb.near_items), static_cast<void*>(tb.near_partial),
                  static_cast<void*>(tb.leaf_chunks), static_cast<void*>(tb.scan_temp)})
    (void)hipFree(p);
  tb.leaf_first = nullptr;
  tb.nodes = nullptr;
  tb.quad = nullptr;
  tb.near_list = tb.near_cnt = tb.far_list = tb.far_cnt = nullptr;
  tb.near_parts = tb.near_offs = nullptr;
  tb.near_items = nullptr;
  tb.near_partial = nullptr;
  tb.leaf_chunks = nullptr;
  tb.scan_temp = nullptr;
  tb.leaf_cap = tb.node_cap = tb.cap_near = tb.cap_far = 0;
}

// buffers indexed by leaf or node, for at least nleaf leaves (grown by a
// quarter beyond, so that the next states rarely need more)
int ensure_leaf_capacity(pw_ctx* c, int nleaf) {
  auto& f = c->ff;
  auto& tb = f.tb;
  if (nleaf <= tb.leaf_cap) return PW_OK;
  PW_HIP(hipStreamSynchronize(c->stream));
  free_leaf_buffers(f);
  const int n = static_cast<int>(c->n);
  const int cap = std::min(n, nleaf + nleaf / 4 + 64);
  const pw::TreeShape sh = pw::tree_shape(n, cap);
  if (sh.nlevel >= pw::kMaxLevels) return fail(PW_ERR_UNSUPPORTED, "too many leaves for the tree");
  const size_t nl = static_cast<size_t>(cap);
  // initial list capacities (PAWAN_FF_LIST_CAP, default 8192; lists grow
  // past them when a tree needs it)
  int cap0 = 8192;
  if (const char* e = std::getenv("PAWAN_FF_LIST_CAP")) cap0 = std::max(1, std::atoi(e));
  const int cap_near = std::min(cap, cap0), cap_far = std::min(sh.nnodes, cap0);
  if (dalloc(&tb.leaf_first, nl + 1) != hipSuccess || dalloc(&tb.nodes, sh.nnodes) != hipSuccess ||
      dalloc(&tb.quad, sh.nnodes) != hipSuccess ||
      dalloc(&tb.near_list, nl * cap_near) != hipSuccess || dalloc(&tb.near_cnt, nl) != hipSuccess ||
      dalloc(&tb.far_list, nl * cap_far) != hipSuccess || dalloc(&tb.far_cnt, nl) != hipSuccess ||
      dalloc(&tb.near_parts, nl) != hipSuccess || dalloc(&tb.near_offs, nl) != hipSuccess ||
      dalloc(&tb.near_items, nl * pw::kNearMaxParts) != hipSuccess ||
      dalloc(&tb.near_partial, nl * pw::kNearMaxParts * pw::kLeafSize) != hipSuccess ||
      dalloc(&tb.leaf_chunks, nl * pw::kLeafChunks) != hipSuccess ||
      pw::near_scan_temp_bytes(cap, &tb.scan_temp_bytes) != hipSuccess ||
      dalloc(reinterpret_cast<unsigned char**>(&tb.scan_temp), tb.scan_temp_bytes) != hipSuccess) {
    (void)hipGetLastError();
    free_leaf_buffers(f);
    return fail(PW_ERR_NOMEM, "tree leaf buffers");
  }
  tb.leaf_cap = cap;
  tb.node_cap = sh.nnodes;
  tb.cap_near = cap_near;
  tb.cap_far = cap_far;
  return PW_OK;
}

// near and far lists that hold at least need_near and need_far entries per
// leaf slot (grown by a quarter beyond; at most every leaf or node)
int ensure_list_capacity(pw_ctx* c, int need_near, int need_far) {
  auto& tb = c->ff.tb;
  if (need_near <= tb.cap_near && need_far <= tb.cap_far) return PW_OK;
  PW_HIP(hipStreamSynchronize(c->stream));
  const int cap_near = std::max(tb.cap_near, std::min(tb.leaf_cap, need_near + need_near / 4 + 64));
  const int cap_far = std::max(tb.cap_far, std::min(tb.node_cap, need_far + need_far / 4 + 64));
  (void)hipFree(tb.near_list);
  (void)hipFree(tb.far_list);
  tb.near_list = tb.far_list = nullptr;
  tb.cap_near = tb.cap_far = 0;
  const size_t nl = static_cast<size_t>(tb.leaf_cap);
  if (dalloc(&tb.near_list, nl * cap_near) != hipSuccess || dalloc(&tb.far_list, nl * cap_far) != hipSuccess) {
    (void)hipGetLastError();
    free_leaf_buffers(c->ff);
    return fail(PW_ERR_NOMEM, "interaction lists");
  }
  tb.cap_near = cap_near;
  tb.cap_far = cap_far;
  return PW_OK;
}

int ensure_ff(pw_ctx* c) {
  PW_TRY(pwi::ensure_built(c));
  if (!c->coupled && c->wn.size() > 1)
    return fail(PW_ERR_UNSUPPORTED, "PW_LEARNED needs a coupled system (one tree over all wakes)");
  auto& f = c->ff;
  if (f.ready) return PW_OK;
  const int n = static_cast<int>(c->n);
  auto& tb = f.tb;
  PW_HIP(pw::tree_sort_temp_bytes(n, &tb.sort_temp_bytes));
  PW_HIP(pw::tree_scan_temp_bytes(n, &tb.scan_n_bytes));
  if (dalloc(&tb.bbox_part, 6 * 1024) != hipSuccess || dalloc(&tb.bbox, 6) != hipSuccess ||
      dalloc(&tb.code, n) != hipSuccess || dalloc(&tb.code_sorted, n) != hipSuccess ||
      dalloc(&tb.idx, n) != hipSuccess || dalloc(&tb.perm, n) != hipSuccess ||
      dalloc(reinterpret_cast<unsigned char**>(&tb.sort_temp), tb.sort_temp_bytes) != hipSuccess ||
      dalloc(&tb.sorted, n) != hipSuccess || dalloc(&tb.cell_start, n) != hipSuccess ||
      dalloc(&tb.cut, n) != hipSuccess || dalloc(&tb.leaf_id, n) != hipSuccess || dalloc(&tb.hist, 16) != hipSuccess ||
      dalloc(reinterpret_cast<unsigned char**>(&tb.scan_n_temp), tb.scan_n_bytes) != hipSuccess ||
      dalloc(&f.d_overflow, 1) != hipSuccess || dalloc(&f.rates_s, n) != hipSuccess ||
      dalloc(&tb.near_total, 1) != hipSuccess || dalloc(&tb.near_queue, 1) != hipSuccess ||
      dalloc(&f.d_dpart, 3 * (static_cast<size_t>(n + 255) / 256 + 1)) != hipSuccess) {
    (void)hipGetLastError();
    pwi::ff_release(c);
    return fail(PW_ERR_NOMEM, "tree buffers");
  }
  tb.overflow = f.d_overflow;
  PW_HIP(pw::launch_fill_int(f.d_overflow, 1, 0, c->stream));
  double smax = 0.0;
  for (double r : c->radius) smax = std::max(smax, r * r);
  f.sigma_max2 = static_cast<float>(smax);
  f.uniform_sigma = !c->radius.empty() &&
                    std::all_of(c->radius.begin(), c->radius.end(), [&](double r) { return r == c->radius[0]; });
  f.ready = true;
  return PW_OK;
}

// mac_only: the lists training samples are drawn from (opening angle and
// the Gaussian core only; the samples' exact sums are those of the singular
// law there, whatever the smoothing kernel)
pw::ListParams list_params(const pw_ctx* c, bool mac_only) {
  const auto& f = c->ff;
  pw::ListParams lp;
  lp.theta = f.theta;
  lp.tol = mac_only ? 0.0f : f.tol;
  lp.err_coef = static_cast<float>(f.use_mlp ? f.coef_mlp : f.coef_md);
  lp.sigma_max2 = f.sigma_max2;
  lp.quad = f.use_mlp ? 0 : 1;  // the network rides on monopole + dipole
  lp.ho_coef = 0.0f;
  lp.ho_gap2 = 0.0f;
  if (c->smoothing == PW_KERNEL_HIGHORDER && !mac_only) {
    // 1 - 4 pi r^3 q = (15/8) x^2 - (35/8) x^3 + ..., x = (sigma/r)^2: at
    // most (15/8) x^2 for r >= sigma, so a node's velocity differs from the
    // singular law by at most (15/8) sigma^4 s0 / (4 pi gap^6)
    const double s4 = static_cast<double>(f.sigma_max2) * f.sigma_max2;
    if (f.tol > 0.0f) lp.ho_coef = static_cast<float>(1.875 * s4 / (4.0 * M_PI));
    else lp.ho_gap2 = 5600.0f;  // (15/8) / gap_bar^4 <= 2^-24
  }
  return lp;
}

// Tree and lists of `state`.  The leaf count is data dependent: it is read
// back once (one stream synchronisation per right-hand side), and the
// leaf-indexed buffers grow when it exceeds them.
int build(pw_ctx* c, const Rec64* state, bool mac_only) {
  auto& f = c->ff;
  const int n = static_cast<int>(c->n);
  PW_HIP(pw::build_leaves(state, n, f.tb, c->stream));
  int nleaf = 0;
  PW_HIP(hipMemcpyAsync(&nleaf, f.tb.hist + 12, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  PW_HIP(hipStreamSynchronize(c->stream));
  if (nleaf < 1 || nleaf > n) return fail(PW_ERR_STATE, "tree: bad leaf count");
  if (nleaf > f.tb.leaf_cap) {
    PW_TRY(ensure_leaf_capacity(c, nleaf));
    PW_HIP(pw::write_leaf_first(n, f.tb, c->stream));
  }
  f.shape = pw::tree_shape(n, nleaf);
  PW_HIP(pw::build_moments(n, f.tb, f.shape, c->stream));
  const pw::ListParams lp = list_params(c, mac_only);
  PW_HIP(pw::build_lists(f.tb, f.shape, lp, c->stream));
  // lists longer than their capacity: grow them and build again (the
  // lengths are those of this tree, so the second build fits)
  int need[2] = {0, 0};
  PW_HIP(hipMemcpyAsync(need, f.tb.hist + pw::kHistNeed, sizeof(need), hipMemcpyDeviceToHost, c->stream));
  PW_HIP(hipStreamSynchronize(c->stream));
  if (need[0] > f.tb.cap_near || need[1] > f.tb.cap_far) {
    PW_TRY(ensure_list_capacity(c, need[0], need[1]));
    PW_HIP(pw::launch_fill_int(f.d_overflow, 1, 0, c->stream));
    PW_HIP(pw::build_lists(f.tb, f.shape, lp, c->stream));
  }
  f.lists_valid = true;
  return PW_OK;
}

int ensure_net(pw_ctx* c) {
  auto& f = c->ff;
  if (f.d_params) return PW_OK;
  if (dalloc(&f.d_params, pw::kMlpParams) != hipSuccess || dalloc(&f.d_m, pw::kMlpParams) != hipSuccess ||
      dalloc(&f.d_v, pw::kMlpParams) != hipSuccess || dalloc(&f.d_grad, pw::kMlpParams) != hipSuccess ||
      dalloc(&f.d_loss, 1) != hipSuccess || dalloc(&f.d_frag, pw::kMlpFragElems) != hipSuccess ||
      dalloc(&f.d_bias, pw::kMlpBiasLen) != hipSuccess) {
    (void)hipGetLastError();
    return fail(PW_ERR_NOMEM, "network buffers");
  }
  return PW_OK;
}

// master parameters (host copy) -> device master, packed BF16 fragments;
// resets the optimizer state
int upload_params(pw_ctx* c) {
  auto& f = c->ff;
  PW_TRY(ensure_net(c));
  std::vector<uint16_t> frag;
  std::vector<float> bias;
  pw::pack_mlp(f.h_params.data(), frag, bias);
  std::vector<float> zero(pw::kMlpParams, 0.0f);
  PW_HIP(hipMemcpy(f.d_params, f.h_params.data(), sizeof(float) * pw::kMlpParams, hipMemcpyHostToDevice));
  PW_HIP(hipMemcpy(f.d_m, zero.data(), sizeof(float) * pw::kMlpParams, hipMemcpyHostToDevice));
  PW_HIP(hipMemcpy(f.d_v, zero.data(), sizeof(float) * pw::kMlpParams, hipMemcpyHostToDevice));
  PW_HIP(hipMemcpy(f.d_frag, frag.data(), sizeof(uint16_t) * frag.size(), hipMemcpyHostToDevice));
  PW_HIP(hipMemcpy(f.d_bias, bias.data(), sizeof(float) * bias.size(), hipMemcpyHostToDevice));
  f.adam_step = 0;
  f.have_weights = true;
  return PW_OK;
}

int repack(pw_ctx* c) {
  auto& f = c->ff;
  std::vector<uint16_t> frag;
  std::vector<float> bias;
  pw::pack_mlp(f.h_params.data(), frag, bias);
  PW_HIP(hipMemcpy(f.d_frag, frag.data(), sizeof(uint16_t) * frag.size(), hipMemcpyHostToDevice));
  PW_HIP(hipMemcpy(f.d_bias, bias.data(), sizeof(float) * bias.size(), hipMemcpyHostToDevice));
  return PW_OK;
}

double quantile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  const size_t k = std::min(v.size() - 1, static_cast<size_t>(q * static_cast<double>(v.size())));
  std::nth_element(v.begin(), v.begin() + static_cast<std::ptrdiff_t>(k), v.end());
  return v[k];
}

}  // namespace

namespace pwi {

void ff_release(pw_ctx* c) {
  auto& f = c->ff;
  auto& tb = f.tb;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto* v : {&f.st_used, &f.st_free})
    for (auto& ev : *v)
      for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  (void)hipFree(tb.bbox_part);
  (void)hipFree(tb.bbox);
  (void)hipFree(tb.code);
  (void)hipFree(tb.code_sorted);
  (void)hipFree(tb.idx);
  (void)hipFree(tb.perm);
  (void)hipFree(tb.sort_temp);
  (void)hipFree(tb.sorted);
  (void)hipFree(tb.nodes);
  (void)hipFree(tb.near_list);
  (void)hipFree(tb.near_cnt);
  (void)hipFree(tb.far_list);
  (void)hipFree(tb.far_cnt);
  (void)hipFree(tb.near_parts);
  (void)hipFree(tb.near_offs);
  (void)hipFree(tb.near_items);
  (void)hipFree(tb.near_total);
  (void)hipFree(tb.near_queue);
  (void)hipFree(tb.near_partial);
  (void)hipFree(tb.leaf_chunks);
  (void)hipFree(tb.scan_temp);
  (void)hipFree(tb.leaf_first);
  (void)hipFree(tb.cell_start);
  (void)hipFree(tb.cut);
  (void)hipFree(tb.leaf_id);
  (void)hipFree(tb.hist);
  (void)hipFree(tb.scan_n_temp);
  (void)hipFree(f.d_overflow);
  (void)hipFree(f.d_dpart);
  (void)hipFree(f.rates_s);
  (void)hipFree(f.d_params);
  (void)hipFree(f.d_m);
  (void)hipFree(f.d_v);
  (void)hipFree(f.d_grad);
  (void)hipFree(f.d_gpart);
  (void)hipFree(f.d_lpart);
  (void)hipFree(f.d_loss);
  (void)hipFree(f.d_frag);
  (void)hipFree(f.d_bias);
  (void)hipFree(f.d_X);
  (void)hipFree(f.d_Y);
  (void)hipFree(f.d_aux);
  const FarField fresh;
  std::vector<float> keep = std::move(f.h_params);
  f = fresh;
  f.h_params = std::move(keep);
}

namespace {

using StageEvents = std::array<hipEvent_t, pwi::FarField::kStages + 1>;

// one right-hand side of the tree path; ev (may be null) gets the stage
// boundaries recorded
int ff_rhs_body(pw_ctx* c, const Rec64* state, Rate64* out, const StageEvents* ev) {
  auto& f = c->ff;
  auto mark = [&](int k) -> hipError_t { return ev ? hipEventRecord((*ev)[k], c->stream) : hipSuccess; };
  PW_HIP(mark(0));
  PW_TRY(build(c, state, false));
  PW_HIP(mark(1));
  const auto& tb = f.tb;
  const int n = static_cast<int>(c->n);
  PW_HIP(pw::build_near_items(tb, f.shape.nleaf, c->stream));
  const bool ho = c->smoothing == PW_KERNEL_HIGHORDER;
  PW_HIP(pw::launch_near_fp32c(tb.sorted, tb.nodes, tb, f.shape.nleaf, c->nu, f.sigma_max2, f.uniform_sigma, ho,
                               c->num_cus * pw::near_kernel_blocks_per_cu(ho, f.uniform_sigma), f.rates_s,
                               c->stream));
  PW_HIP(mark(2));
  PW_HIP(pw::launch_far_md(tb.sorted, tb.nodes, f.use_mlp ? nullptr : tb.quad, tb.far_list, tb.far_cnt,
                           tb.cap_far, f.shape.nleaf, f.rates_s, c->stream));
  PW_HIP(mark(3));
  if (f.use_mlp) {
    pw::MlpWeights W{f.d_frag, f.d_bias};
    PW_HIP(pw::launch_far_mlp(tb.sorted, tb.nodes, tb.far_list, tb.far_cnt, tb.cap_far, f.shape.nleaf, W,
                              f.rates_s, c->stream));
  }
  PW_HIP(mark(4));
  if (f.conserve) PW_HIP(pw::launch_remove_mean_dalpha(f.rates_s, n, f.d_dpart, c->stream));
  PW_HIP(pw::launch_scatter_rates(f.rates_s, tb.perm, n, out, c->stream));
  PW_HIP(mark(5));
  return PW_OK;
}

}  // namespace

int ff_rhs(pw_ctx* c, const Rec64* state, Rate64* out) {
  PW_TRY(ensure_ff(c));
  auto& f = c->ff;
  if (f.use_mlp && !f.have_weights)
    return fail(PW_ERR_STATE, "PW_LEARNED with the network on needs weights (pw_ff_train or pw_ff_set_weights), "
                              "or pw_ff_set_params(..., use_mlp = 0)");
  if (!c->profiling) return ff_rhs_body(c, state, out, nullptr);
  // stage events while the context profiles (recycled; a failed right-hand
  // side returns its events unread)
  StageEvents ev{};
  if (f.st_free.empty()) {
    for (auto& e : ev) PW_HIP(hipEventCreate(&e));
  } else {
    ev = f.st_free.back();
    f.st_free.pop_back();
  }
  const int rc = ff_rhs_body(c, state, out, &ev);
  (rc == PW_OK ? f.st_used : f.st_free).push_back(ev);
  return rc;
}

int ff_collect_stages(pw_ctx* c) {
  auto& f = c->ff;
  if (f.st_used.empty()) return PW_OK;
  PW_HIP(hipStreamSynchronize(c->stream));
  for (auto& ev : f.st_used) {
    for (int k = 0; k < FarField::kStages; ++k) {
      float ms = 0.0f;
      PW_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
      f.stage_ms[k] += ms;
    }
    f.stage_rhs += 1;
    f.st_free.push_back(ev);
  }
  f.st_used.clear();
  return PW_OK;
}

void ff_reset_stages(pw_ctx* c) {
  auto& f = c->ff;
  for (double& v : f.stage_ms) v = 0.0;
  f.stage_rhs = 0;
}

int ff_check(pw_ctx* c) {
  auto& f = c->ff;
  if (!f.ready) return PW_OK;
  int over = 0;
  PW_HIP(hipMemcpyAsync(&over, f.d_overflow, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  PW_HIP(hipStreamSynchronize(c->stream));
  if (over) {
    PW_HIP(pw::launch_fill_int(f.d_overflow, 1, 0, c->stream));
    return fail(PW_ERR_STATE, "interaction lists exceeded their capacity (near " + std::to_string(f.tb.cap_near) +
                                  ", far " + std::to_string(f.tb.cap_far) + "); use a smaller theta or tolerance");
  }
  return PW_OK;
}

}  // namespace pwi

extern "C" {

int pw_ff_set_params(pw_ctx* c, double theta, double tol, int use_mlp) {
  PW_CTX(c);
  if (!(theta > 0.0 && theta < 1.0)) return fail(PW_ERR_ARG, "theta must be in (0, 1)");
  if (!std::isfinite(tol)) return fail(PW_ERR_ARG, "tol must be finite");
  c->ff.theta = static_cast<float>(theta);
  c->ff.tol = static_cast<float>(std::max(0.0, tol));
  c->ff.use_mlp = use_mlp != 0;
  return PW_OK;
}

int pw_ff_get_params(pw_ctx* c, double* theta, double* tol, int* use_mlp) {
  PW_CTX(c);
  if (theta) *theta = c->ff.theta;
  if (tol) *tol = c->ff.tol;
  if (use_mlp) *use_mlp = c->ff.use_mlp ? 1 : 0;
  return PW_OK;
}

int pw_ff_set_conserve(pw_ctx* c, int enable) {
  PW_CTX(c);
  c->ff.conserve = enable != 0;
  return PW_OK;
}

int pw_ff_get_conserve(pw_ctx* c, int* enable) {
  PW_CTX(c);
  if (!enable) return fail(PW_ERR_ARG, "null output");
  *enable = c->ff.conserve ? 1 : 0;
  return PW_OK;
}

int pw_ff_set_err_coef(pw_ctx* c, double coef_md, double coef_mlp) {
  PW_CTX(c);
  if (!(coef_md >= 0.0) || !(coef_mlp >= 0.0) || !std::isfinite(coef_md) || !std::isfinite(coef_mlp))
    return fail(PW_ERR_ARG, "coefficients must be finite and >= 0");
  c->ff.coef_md = coef_md;
  c->ff.coef_mlp = coef_mlp;
  return PW_OK;
}

int pw_ff_get_err_coef(pw_ctx* c, double* coef_md, double* coef_mlp) {
  PW_CTX(c);
  if (coef_md) *coef_md = c->ff.coef_md;
  if (coef_mlp) *coef_mlp = c->ff.coef_mlp;
  return PW_OK;
}

int pw_ff_build(pw_ctx* c) {
  PW_CTX(c);
  PW_TRY(ensure_ff(c));
  PW_TRY(build(c, c->y, false));
  return pwi::ff_check(c);
}

int pw_ff_info(pw_ctx* c, int64_t* out, int64_t len) {
  PW_CTX(c);
  if (!out || len < 1) return fail(PW_ERR_ARG, "null output");
  PW_TRY(ensure_ff(c));
  auto& f = c->ff;
  std::vector<int64_t> v(14, 0);
  v[0] = c->n;
  v[1] = f.shape.nleaf;
  v[2] = f.shape.nnodes;
  v[3] = f.shape.nlevel;
  v[4] = f.tb.cap_near;
  v[5] = f.tb.cap_far;
  v[6] = f.ns;
  if (f.lists_valid) {
    long long* d = nullptr;
    if (dalloc(&d, 3 * static_cast<size_t>(f.shape.nleaf)) != hipSuccess) return fail(PW_ERR_NOMEM, "stats");
    std::vector<long long> h(3 * static_cast<size_t>(f.shape.nleaf));
    std::vector<int> nc(f.shape.nleaf), fc(f.shape.nleaf);
    hipError_t e = pw::launch_list_stats(f.tb, f.shape, f.sigma_max2, d, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(h.data(), d, h.size() * sizeof(long long), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(nc.data(), f.tb.near_cnt, nc.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(fc.data(), f.tb.far_cnt, fc.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    int over = 0;
    if (e == hipSuccess)
      e = hipMemcpyAsync(&over, f.d_overflow, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(PW_ERR_HIP, std::string("list stats: ") + hipGetErrorString(e));
    for (int T = 0; T < f.shape.nleaf; ++T) {
      v[7] += h[3 * T];
      v[8] += h[3 * T + 1];
      v[12] += h[3 * T + 2];
      v[10] = std::max<int64_t>(v[10], nc[T]);
      v[11] = std::max<int64_t>(v[11], fc[T]);
      v[13] += fc[T];
    }
    v[9] = over;
  }
  for (int64_t i = 0; i < std::min<int64_t>(len, 14); ++i) out[i] = v[i];
  return PW_OK;
}

int pw_ff_get_timing(pw_ctx* c, double* out, int64_t len) {
  PW_CTX(c);
  if (!out || len < 1) return fail(PW_ERR_ARG, "null output");
  PW_TRY(pwi::ff_collect_stages(c));
  const auto& f = c->ff;
  double v[1 + pwi::FarField::kStages];
  v[0] = static_cast<double>(f.stage_rhs);
  for (int k = 0; k < pwi::FarField::kStages; ++k) v[1 + k] = f.stage_ms[k];
  for (int64_t i = 0; i < std::min<int64_t>(len, 1 + pwi::FarField::kStages); ++i) out[i] = v[i];
  return PW_OK;
}

int pw_ff_export(pw_ctx* c, int what, void* out, int64_t bytes) {
  PW_CTX(c);
  if (!out) return fail(PW_ERR_ARG, "null output");
  PW_TRY(ensure_ff(c));
  auto& f = c->ff;
  const auto& tb = f.tb;
  const int64_t nl = f.shape.nleaf;
  auto copy = [&](const void* src, int64_t need) -> int {
    if (bytes != need) return fail(PW_ERR_ARG, "export needs exactly " + std::to_string(need) + " bytes");
    PW_HIP(hipMemcpyAsync(out, src, static_cast<size_t>(need), hipMemcpyDeviceToHost, c->stream));
    PW_HIP(hipStreamSynchronize(c->stream));
    return PW_OK;
  };
  switch (what) {
    case PW_FF_PERM: return copy(tb.perm, 4 * c->n);
    case PW_FF_NEAR_CNT: return copy(tb.near_cnt, 4 * nl);
    case PW_FF_NEAR: return copy(tb.near_list, 4 * nl * tb.cap_near);
    case PW_FF_FAR_CNT: return copy(tb.far_cnt, 4 * nl);
    case PW_FF_FAR: return copy(tb.far_list, 4 * nl * tb.cap_far);
    case PW_FF_SAMPLES_X: return copy(f.d_X, 4 * 16 * f.ns);
    case PW_FF_SAMPLES_Y: return copy(f.d_Y, 4 * 12 * f.ns);
    case PW_FF_SAMPLES_AUX: return copy(f.d_aux, 8 * pw::kSampleAux * f.ns);
    case PW_FF_NODES: {
      const int64_t nn = f.shape.nnodes;
      if (bytes != 8 * 24 * nn) return fail(PW_ERR_ARG, "export needs exactly " + std::to_string(8 * 24 * nn) + " bytes");
      std::vector<pw::NodeRec> h(static_cast<size_t>(nn));
      PW_HIP(hipMemcpyAsync(h.data(), tb.nodes, h.size() * sizeof(pw::NodeRec), hipMemcpyDeviceToHost, c->stream));
      PW_HIP(hipStreamSynchronize(c->stream));
      double* o = static_cast<double*>(out);
      for (int64_t i = 0; i < nn; ++i, o += 24) {
        const pw::NodeRec& q = h[static_cast<size_t>(i)];
        o[0] = q.cx; o[1] = q.cy; o[2] = q.cz;
        o[3] = q.r; o[4] = q.s0; o[5] = q.s2; o[6] = q.sg;
        for (int k = 0; k < 3; ++k) o[7 + k] = q.M[k];
        for (int k = 0; k < 9; ++k) o[10 + k] = q.D[k];
        o[19] = q.first; o[20] = q.last; o[21] = q.level; o[22] = q.child;
        o[23] = 0.0;
      }
      return PW_OK;
    }
    case PW_FF_QUAD: {
      const int64_t nn = f.shape.nnodes;
      if (bytes != 8 * 20 * nn) return fail(PW_ERR_ARG, "export needs exactly " + std::to_string(8 * 20 * nn) + " bytes");
      std::vector<pw::NodeRec> h(static_cast<size_t>(nn));
      std::vector<pw::NodeQuad> hq(static_cast<size_t>(nn));
      PW_HIP(hipMemcpyAsync(h.data(), tb.nodes, h.size() * sizeof(pw::NodeRec), hipMemcpyDeviceToHost, c->stream));
      PW_HIP(hipMemcpyAsync(hq.data(), tb.quad, hq.size() * sizeof(pw::NodeQuad), hipMemcpyDeviceToHost, c->stream));
      PW_HIP(hipStreamSynchronize(c->stream));
      double* o = static_cast<double*>(out);
      for (int64_t i = 0; i < nn; ++i, o += 20) {
        for (int k = 0; k < pw::kQuad; ++k) o[k] = hq[static_cast<size_t>(i)].Q[k];
        o[18] = h[static_cast<size_t>(i)].s1;
        o[19] = h[static_cast<size_t>(i)].s3;
      }
      return PW_OK;
    }
    default:
      return fail(PW_ERR_ARG, "unknown export");
  }
}

int pw_ff_sample(pw_ctx* c, int64_t nsamples, uint64_t seed) {
  PW_CTX(c);
  if (nsamples <= 0 || nsamples > (int64_t(1) << 26)) return fail(PW_ERR_ARG, "nsamples must be in [1, 2^26]");
  PW_TRY(ensure_ff(c));
  auto& f = c->ff;
  PW_TRY(build(c, c->y, true));  // opening angle only: every candidate far node
  f.lists_valid = false;         // the next right-hand side rebuilds with the tolerance
  const int nl = f.shape.nleaf;
  std::vector<int> fc(nl), lf(nl + 1);
  PW_HIP(hipMemcpyAsync(fc.data(), f.tb.far_cnt, sizeof(int) * nl, hipMemcpyDeviceToHost, c->stream));
  PW_HIP(hipMemcpyAsync(lf.data(), f.tb.leaf_first, sizeof(int) * (nl + 1), hipMemcpyDeviceToHost, c->stream));
  PW_HIP(hipStreamSynchronize(c->stream));
  PW_TRY(pwi::ff_check(c));
  std::vector<int> with;
  for (int T = 0; T < nl; ++T)
    if (fc[T] > 0) with.push_back(T);
  if (with.empty()) return fail(PW_ERR_STATE, "no far interactions at this opening angle");
  std::vector<int> spec(3 * static_cast<size_t>(nsamples));
  uint64_t s = seed;
  for (int64_t k = 0; k < nsamples; ++k) {
    const int T = with[splitmix(s) % with.size()];
    const int slot = static_cast<int>(splitmix(s) % static_cast<uint64_t>(fc[T]));
    const int size = lf[T + 1] - lf[T];
    const int t = lf[T] + static_cast<int>(splitmix(s) % static_cast<uint64_t>(size));
    spec[3 * k] = T;
    spec[3 * k + 1] = slot;
    spec[3 * k + 2] = t;
  }
  // grow the sample set
  const int64_t need = f.ns + nsamples;
  if (need > f.ns_cap) {
    const int64_t cap = std::max<int64_t>(need, 2 * f.ns_cap);
    float *X = nullptr, *Y = nullptr;
    double* A = nullptr;
    if (dalloc(&X, 16 * cap) != hipSuccess || dalloc(&Y, 12 * cap) != hipSuccess ||
        dalloc(&A, pw::kSampleAux * cap) != hipSuccess) {
      (void)hipFree(X);
      (void)hipFree(Y);
      (void)hipFree(A);
      (void)hipGetLastError();
      return fail(PW_ERR_NOMEM, "sample buffers");
    }
    if (f.ns > 0) {
      PW_HIP(hipMemcpyAsync(X, f.d_X, sizeof(float) * 16 * f.ns, hipMemcpyDeviceToDevice, c->stream));
      PW_HIP(hipMemcpyAsync(Y, f.d_Y, sizeof(float) * 12 * f.ns, hipMemcpyDeviceToDevice, c->stream));
      PW_HIP(hipMemcpyAsync(A, f.d_aux, sizeof(double) * pw::kSampleAux * f.ns, hipMemcpyDeviceToDevice, c->stream));
      PW_HIP(hipStreamSynchronize(c->stream));
    }
    (void)hipFree(f.d_X);
    (void)hipFree(f.d_Y);
    (void)hipFree(f.d_aux);
    f.d_X = X;
    f.d_Y = Y;
    f.d_aux = A;
    f.ns_cap = cap;
  }
  int* d_spec = nullptr;
  if (dalloc(&d_spec, spec.size()) != hipSuccess) return fail(PW_ERR_NOMEM, "sample spec");
  hipError_t e = hipMemcpyAsync(d_spec, spec.data(), sizeof(int) * spec.size(), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = pw::launch_ff_samples(f.tb.sorted, f.tb.nodes, f.tb.quad, f.tb.far_list, f.tb.cap_far, d_spec,
                              static_cast<int>(nsamples), f.d_X + 16 * f.ns, f.d_Y + 12 * f.ns,
                              f.d_aux + pw::kSampleAux * f.ns, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d_spec);
  if (e != hipSuccess) return fail(PW_ERR_HIP, std::string("sample generation: ") + hipGetErrorString(e));
  f.ns += nsamples;
  return PW_OK;
}

int pw_ff_clear_samples(pw_ctx* c) {
  PW_CTX(c);
  c->ff.ns = 0;
  return PW_OK;
}

int pw_ff_init_weights(pw_ctx* c, uint64_t seed) {
  PW_CTX(c);
  c->ff.h_params.assign(pw::kMlpParams, 0.0f);
  pw::init_mlp(seed, c->ff.h_params.data());
  return upload_params(c);
}

int pw_ff_set_weights(pw_ctx* c, const float* params, int64_t n) {
  PW_CTX(c);
  if (!params || n != pw::kMlpParams) return fail(PW_ERR_ARG, "need 36748 parameters");
  for (int64_t i = 0; i < n; ++i)
    if (!std::isfinite(params[i])) return fail(PW_ERR_ARG, "non-finite parameter");
  c->ff.h_params.assign(params, params + n);
  return upload_params(c);
}

int pw_ff_get_weights(pw_ctx* c, float* params, int64_t n) {
  PW_CTX(c);
  if (!params || n != pw::kMlpParams) return fail(PW_ERR_ARG, "need room for 36748 parameters");
  if (!c->ff.have_weights) return fail(PW_ERR_STATE, "no weights");
  PW_HIP(hipMemcpy(params, c->ff.d_params, sizeof(float) * n, hipMemcpyDeviceToHost));
  return PW_OK;
}

int pw_ff_train(pw_ctx* c, int64_t steps, int batch, double lr, uint64_t seed, double* report,
                int64_t report_len) {
  PW_CTX(c);
  auto& f = c->ff;
  if (steps < 0 || batch < 32 || batch > 65536 || !(lr > 0.0))
    return fail(PW_ERR_ARG, "need steps >= 0, 32 <= batch <= 65536, lr > 0");
  const int64_t nval = f.ns / 10;
  const int64_t ntrain = f.ns - nval;
  if (ntrain < batch || nval < 16) return fail(PW_ERR_STATE, "too few samples (pw_ff_sample first)");
  if (!f.have_weights) PW_TRY(pw_ff_init_weights(c, seed ^ 0x5bd1e995ull));
  const auto t0 = std::chrono::steady_clock::now();
  const int blocks = pw::train_blocks(batch);
  if (f.gpart_blocks < blocks) {
    (void)hipFree(f.d_gpart);
    (void)hipFree(f.d_lpart);
    f.d_gpart = nullptr;
    f.d_lpart = nullptr;
    f.gpart_blocks = 0;
    if (dalloc(&f.d_gpart, static_cast<size_t>(blocks) * pw::kMlpParams) != hipSuccess ||
        dalloc(&f.d_lpart, blocks) != hipSuccess)
      return fail(PW_ERR_NOMEM, "gradient buffers");
    f.gpart_blocks = blocks;
  }
  int* d_idx = nullptr;
  if (dalloc(&d_idx, ntrain) != hipSuccess) return fail(PW_ERR_NOMEM, "index buffer");
  std::vector<int> order(static_cast<size_t>(ntrain));
  std::iota(order.begin(), order.end(), 0);
  uint64_t s = seed;
  int64_t pos = ntrain;  // force a shuffle at the start
  float last_loss = 0.0f;
  int rc = PW_OK;
  for (int64_t it = 0; it < steps && rc == PW_OK; ++it) {
    if (pos + batch > ntrain) {
      for (int64_t i = ntrain - 1; i > 0; --i)
        std::swap(order[i], order[splitmix(s) % static_cast<uint64_t>(i + 1)]);
      if (hipMemcpyAsync(d_idx, order.data(), sizeof(int) * ntrain, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
        rc = fail(PW_ERR_HIP, "index upload");
        break;
      }
      pos = 0;
    }
    const double frac = steps > 1 ? static_cast<double>(it) / static_cast<double>(steps - 1) : 1.0;
    const float lr_t = static_cast<float>(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(M_PI * frac))));
    hipError_t e = pw::launch_mlp_grad(f.d_params, f.d_X, f.d_Y, d_idx + pos, batch, f.d_gpart, f.d_lpart,
                                       f.d_grad, f.d_loss, c->stream);
    if (e == hipSuccess)
      e = pw::launch_adam(f.d_params, f.d_grad, f.d_m, f.d_v, lr_t, 0.9f, 0.999f, 1e-8f, ++f.adam_step, c->stream);
    if (e != hipSuccess) rc = fail(PW_ERR_HIP, std::string("training step: ") + hipGetErrorString(e));
    pos += batch;
  }
  if (rc == PW_OK && steps > 0) {
    if (hipMemcpyAsync(&last_loss, f.d_loss, sizeof(float), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      rc = fail(PW_ERR_HIP, "loss readback");
  }
  (void)hipFree(d_idx);
  PW_TRY(rc);
  // master -> host copy -> BF16 fragments
  f.h_params.resize(pw::kMlpParams);
  PW_HIP(hipMemcpy(f.h_params.data(), f.d_params, sizeof(float) * pw::kMlpParams, hipMemcpyDeviceToHost));
  for (float p : f.h_params)
    if (!std::isfinite(p)) return fail(PW_ERR_NONFINITE, "training diverged (non-finite weight); lower lr");
  PW_TRY(repack(c));
  // held-out evaluation with the inference kernel
  float* d_hat = nullptr;
  if (dalloc(&d_hat, 12 * nval) != hipSuccess) return fail(PW_ERR_NOMEM, "held-out outputs");
  pw::MlpWeights W{f.d_frag, f.d_bias};
  std::vector<float> hat(12 * nval), y(12 * nval);
  std::vector<double> aux(pw::kSampleAux * nval);
  hipError_t e = pw::launch_mlp_forward(f.d_X + 16 * ntrain, static_cast<int>(nval), d_hat, W, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(hat.data(), d_hat, sizeof(float) * hat.size(), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(y.data(), f.d_Y + 12 * ntrain, sizeof(float) * y.size(), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(aux.data(), f.d_aux + pw::kSampleAux * ntrain, sizeof(double) * aux.size(),
                       hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d_hat);
  if (e != hipSuccess) return fail(PW_ERR_HIP, std::string("held-out evaluation: ") + hipGetErrorString(e));
  double loss_net = 0, loss_zero = 0, eu_md = 0, eu_net = 0, eg_md = 0, eg_net = 0;
  std::vector<double> r_md, r_net;
  r_md.reserve(nval);
  r_net.reserve(nval);
  for (int64_t k = 0; k < nval; ++k) {
    const float* yk = &y[12 * k];
    const float* hk = &hat[12 * k];
    const double* ak = &aux[pw::kSampleAux * k];
    double um = 0, un = 0, gm = 0, gn = 0;
    for (int o = 0; o < 12; ++o) {
      const double dn = static_cast<double>(hk[o]) - yk[o];
      loss_net += dn * dn;
      loss_zero += static_cast<double>(yk[o]) * yk[o];
      if (o < 3) {
        um += static_cast<double>(yk[o]) * yk[o];
        un += dn * dn;
      } else {
        gm += static_cast<double>(yk[o]) * yk[o];
        gn += dn * dn;
      }
    }
    // true-unit errors: scaled residual * scale / (4 pi)
    const double su = ak[0] * kQ, sg = ak[1] * kQ;
    eu_md += um * su * su;
    eu_net += un * su * su;
    eg_md += gm * sg * sg;
    eg_net += gn * sg * sg;
    // the analytic path (quadrupole order) against s3 / d^5, the network
    // path (monopole + dipole + network) against s2 / d^4
    const double d = ak[2];
    const double d4 = d * d * d * d;
    const double bound_q = ak[11] / (d4 * d), bound_d = ak[3] / d4;
    if (bound_q > 0) r_md.push_back(ak[10] * kQ / bound_q);
    if (bound_d > 0) r_net.push_back(std::sqrt(un) * su / bound_d);
  }
  f.coef_md = quantile(r_md, 0.99);
  f.coef_mlp = quantile(r_net, 0.99);
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  const double rep[10] = {static_cast<double>(last_loss),
                          loss_net / static_cast<double>(nval),
                          loss_zero / static_cast<double>(nval),
                          eu_md > 0 ? std::sqrt(eu_net / eu_md) : 0.0,
                          eg_md > 0 ? std::sqrt(eg_net / eg_md) : 0.0,
                          f.coef_md,
                          f.coef_mlp,
                          static_cast<double>(steps),
                          static_cast<double>(nval),
                          ms};
  if (report)
    for (int64_t i = 0; i < std::min<int64_t>(report_len, 10); ++i) report[i] = rep[i];
  return PW_OK;
}

int pw_ff_mlp_forward(pw_ctx* c, const float* X, int64_t n, float* Y) {
  PW_CTX(c);
  if (!X || !Y || n <= 0 || n > (int64_t(1) << 28)) return fail(PW_ERR_ARG, "bad arrays");
  if (!c->ff.have_weights) return fail(PW_ERR_STATE, "no weights");
  float *dX = nullptr, *dY = nullptr;
  if (dalloc(&dX, 16 * n) != hipSuccess || dalloc(&dY, 12 * n) != hipSuccess) {
    (void)hipFree(dX);
    (void)hipGetLastError();
    return fail(PW_ERR_NOMEM, "forward buffers");
  }
  pw::MlpWeights W{c->ff.d_frag, c->ff.d_bias};
  hipError_t e = hipMemcpyAsync(dX, X, sizeof(float) * 16 * n, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = pw::launch_mlp_forward(dX, static_cast<int>(n), dY, W, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(Y, dY, sizeof(float) * 12 * n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dX);
  (void)hipFree(dY);
  if (e != hipSuccess) return fail(PW_ERR_HIP, std::string("mlp forward: ") + hipGetErrorString(e));
  return PW_OK;
}

int pw_ff_mlp_grad(pw_ctx* c, const float* X, const float* Y, int64_t n, float* grad, double* loss) {
  PW_CTX(c);
  if (!X || !Y || !grad || n <= 0 || n > (int64_t(1) << 22)) return fail(PW_ERR_ARG, "bad arrays");
  if (!c->ff.have_weights) return fail(PW_ERR_STATE, "no weights");
  const int blocks = pw::train_blocks(static_cast<int>(n));
  float *dX = nullptr, *dY = nullptr, *gp = nullptr, *lp = nullptr, *g = nullptr, *l = nullptr;
  int* idx = nullptr;
  std::vector<int> h_idx(static_cast<size_t>(n));
  std::iota(h_idx.begin(), h_idx.end(), 0);
  hipError_t e = hipSuccess;
  if (dalloc(&dX, 16 * n) != hipSuccess || dalloc(&dY, 12 * n) != hipSuccess ||
      dalloc(&gp, static_cast<size_t>(blocks) * pw::kMlpParams) != hipSuccess || dalloc(&lp, blocks) != hipSuccess ||
      dalloc(&g, pw::kMlpParams) != hipSuccess || dalloc(&l, 1) != hipSuccess || dalloc(&idx, n) != hipSuccess)
    e = hipErrorOutOfMemory;
  float lh = 0.0f;
  if (e == hipSuccess) e = hipMemcpyAsync(dX, X, sizeof(float) * 16 * n, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dY, Y, sizeof(float) * 12 * n, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(idx, h_idx.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = pw::launch_mlp_grad(c->ff.d_params, dX, dY, idx, static_cast<int>(n), gp, lp, g, l, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(grad, g, sizeof(float) * pw::kMlpParams, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&lh, l, sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  for (void* p : {static_cast<void*>(dX), static_cast<void*>(dY), static_cast<void*>(gp), static_cast<void*>(lp),
                  static_cast<void*>(g), static_cast<void*>(l), static_cast<void*>(idx)})
    (void)hipFree(p);
  (void)hipGetLastError();
  if (e != hipSuccess) return fail(PW_ERR_HIP, std::string("mlp grad: ") + hipGetErrorString(e));
  if (loss) *loss = lh;
  return PW_OK;
}

}  // extern "C"
