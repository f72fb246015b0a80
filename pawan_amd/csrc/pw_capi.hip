// pw_capi.hip -- solver context and the extern "C" entry points declared in
// include/pawan_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <string>
#include <utility>
#include <vector>

#include "pw_ctx.h"

using pw::Rate64;
using pw::Rec64;
using pw::WorkItem;
using pwi::dalloc;
using pwi::fail;
using pwi::kKindLearned;
using pwi::kTargetsPerItem;

namespace pwi {

namespace {
thread_local std::string g_error;
}  // namespace

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

}  // namespace pwi

// Staging of pw_velocity_at_noalloc: the points, then their velocities, for
// at most kNoAllocPoints points (3 MiB).  Every GPU the library runs on holds
// its own copy until the process ends; host code reaches it only through the
// per-device address that hipGetSymbolAddress gives (noalloc_staging).
constexpr int64_t kNoAllocPoints = 65536;
static __device__ __align__(256) double scratch[2 * 3 * kNoAllocPoints];

namespace {

// Device address of scratch on each device, resolved by pw_create with the
// device current, once per device.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int resolve_noalloc_staging(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(PW_ERR_UNSUPPORTED, "device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return PW_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(PW_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = static_cast<double*>(p);
  return PW_OK;
}

double* noalloc_staging(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

// Device buffers of the context.  Captured step graphs hold kernel nodes
// whose arguments point into these buffers, so they are destroyed first,
// after the stream has drained.
void free_buffers(pw_ctx* c) {
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.second);
  c->graphs.clear();
  (void)hipFree(c->y);
  (void)hipFree(c->ys);
  for (auto& p : c->k) (void)hipFree(p);
  (void)hipFree(c->d_radius);
  (void)hipFree(c->pair_items);
  (void)hipFree(c->lin_items);
  (void)hipFree(c->d_lin);
  (void)hipFree(c->d_quad);
  (void)hipFree(c->d_flag);
  for (auto& k : c->ks) (void)hipFree(k.d_items);
  (void)hipFree(c->partial);
  (void)hipFree(c->d_queue);
  (void)hipFree(c->d_chunks);
  (void)hipFree(c->d_src32);
  (void)hipFree(c->d_src64);
  c->d_chunks = nullptr;
  c->d_src32 = nullptr;
  c->d_src64 = nullptr;
  c->y = c->ys = nullptr;
  for (auto& p : c->k) p = nullptr;
  c->rates = nullptr;
  c->d_radius = nullptr;
  c->pair_items = c->lin_items = nullptr;
  c->d_lin = c->d_quad = nullptr;
  c->d_flag = c->d_queue = nullptr;
  for (auto& k : c->ks) {
    k.d_items = nullptr;
    k.nitems = 0;
    k.built_parts = 0;
  }
  c->partial = nullptr;
  c->partial_parts = 0;
  pwi::ff_release(c);
}

// Everything the context holds on the device: buffers, events, stream.
void release(pw_ctx* c) {
  free_buffers(c);
  for (auto& e : c->ev_used) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  for (auto& e : c->ev_free) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

// index of the source chunk holding particle s (chunk_first is ascending)
int32_t chunk_of(const pw_ctx* c, int64_t s) {
  const auto it = std::upper_bound(c->chunk_first.begin(), c->chunk_first.end(), static_cast<int32_t>(s));
  return static_cast<int32_t>(it - c->chunk_first.begin()) - 1;
}

// Work items: blocks of up to 256 consecutive targets of one system, each
// against all sources of that system.  For independent wakes the largest
// systems come first (longest-processing-time order), so the tail of the
// grid is made of short blocks.
void build_items(pw_ctx* c) {
  c->h_pair_items.clear();
  c->pair_sys.clear();
  c->h_lin_items.clear();
  c->lin_wake.clear();
  const int nw = static_cast<int>(c->wn.size());
  struct Sys {
    int64_t a, b;
    int id;
  };
  std::vector<Sys> sys;
  if (c->coupled)
    sys.push_back({0, c->n, 0});
  else
    for (int w = 0; w < nw; ++w) sys.push_back({c->woff[w], c->woff[w + 1], w});
  std::stable_sort(sys.begin(), sys.end(),
                   [](const Sys& p, const Sys& q) { return (p.b - p.a) > (q.b - q.a); });
  c->systems.clear();
  for (const Sys& s : sys) c->systems.push_back({s.a, s.b, s.id});
  for (const Sys& s : sys)
    for (int64_t t0 = s.a; t0 < s.b; t0 += kTargetsPerItem) {
      c->h_pair_items.push_back({static_cast<int32_t>(t0),
                                 static_cast<int32_t>(std::min<int64_t>(t0 + kTargetsPerItem, s.b)),
                                 static_cast<int32_t>(s.a), static_cast<int32_t>(s.b), -1, 0, 0, 0});
      c->pair_sys.push_back(s.id);
    }
  for (int w = 0; w < nw; ++w)
    for (int64_t t0 = c->woff[w]; t0 < c->woff[w + 1]; t0 += kTargetsPerItem) {
      c->h_lin_items.push_back(
          {static_cast<int32_t>(t0),
           static_cast<int32_t>(std::min<int64_t>(t0 + kTargetsPerItem, c->woff[w + 1])), w, 0,
           0, 0, 0, 0});
      c->lin_wake.push_back(w);
    }
  c->npair = static_cast<int>(c->h_pair_items.size());
  c->nlin = static_cast<int>(c->h_lin_items.size());
  // source chunks of the scalar-fed FP32C kernel: runs of up to 64
  // particles of one system, in particle order
  c->chunk_first.clear();
  if (c->coupled) {
    for (int64_t i = 0; i < c->n; i += 64) c->chunk_first.push_back(static_cast<int32_t>(i));
  } else {
    for (int w = 0; w < nw; ++w)
      for (int64_t i = c->woff[w]; i < c->woff[w + 1]; i += 64) c->chunk_first.push_back(static_cast<int32_t>(i));
  }
  c->nchunks = static_cast<int>(c->chunk_first.size());
  for (auto& it : c->h_pair_items) it.chunk0 = chunk_of(c, it.s0);
}

int ensure_built(pw_ctx* c) {
  if (c->built) return PW_OK;
  if (c->n <= 0) return fail(PW_ERR_STATE, "context has no particles (add a wake first)");
  if (c->n > INT32_MAX / 2) return fail(PW_ERR_UNSUPPORTED, "too many particles");
  PW_HIP(hipSetDevice(c->device));
  build_items(c);
  const size_t n = static_cast<size_t>(c->n);
  if (dalloc(&c->y, n) != hipSuccess || dalloc(&c->ys, n) != hipSuccess ||
      dalloc(&c->k[0], n) != hipSuccess || dalloc(&c->k[1], n) != hipSuccess ||
      dalloc(&c->k[2], n) != hipSuccess || dalloc(&c->k[3], n) != hipSuccess ||
      dalloc(&c->d_radius, n) != hipSuccess ||
      dalloc(&c->pair_items, c->h_pair_items.size()) != hipSuccess ||
      dalloc(&c->lin_items, c->h_lin_items.size()) != hipSuccess ||
      dalloc(&c->d_lin, 11 * c->h_lin_items.size()) != hipSuccess ||
      dalloc(&c->d_quad, pw::kDiagQuad * c->h_pair_items.size()) != hipSuccess ||
      dalloc(&c->d_flag, 1) != hipSuccess || dalloc(&c->d_queue, 1) != hipSuccess ||
      dalloc(&c->d_chunks, static_cast<size_t>(c->nchunks)) != hipSuccess ||
      dalloc(&c->d_src32, n) != hipSuccess) {
    free_buffers(c);  // the stream stays: the context remains usable
    (void)hipGetLastError();
    return fail(PW_ERR_NOMEM, "device allocation failed");
  }
  PW_HIP(hipMemcpyAsync(c->y, c->pending.data(), n * sizeof(Rec64), hipMemcpyHostToDevice,
                        c->stream));
  PW_HIP(hipMemcpyAsync(c->d_radius, c->radius.data(), n * sizeof(double),
                        hipMemcpyHostToDevice, c->stream));
  PW_HIP(hipMemcpyAsync(c->pair_items, c->h_pair_items.data(),
                        c->h_pair_items.size() * sizeof(WorkItem), hipMemcpyHostToDevice,
                        c->stream));
  {
    // chunk extents now; their spheres are rebuilt before every launch
    std::vector<pw::Chunk32> ch(static_cast<size_t>(c->nchunks));
    for (int k = 0; k < c->nchunks; ++k) {
      std::memset(&ch[k], 0, sizeof(pw::Chunk32));
      ch[k].first = c->chunk_first[k];
      const int64_t end = k + 1 < c->nchunks ? c->chunk_first[k + 1] : c->n;
      ch[k].count = static_cast<int32_t>(std::min<int64_t>(64, end - c->chunk_first[k]));
    }
    PW_HIP(hipMemcpy(c->d_chunks, ch.data(), ch.size() * sizeof(pw::Chunk32), hipMemcpyHostToDevice));
  }
  PW_HIP(hipMemcpyAsync(c->lin_items, c->h_lin_items.data(),
                        c->h_lin_items.size() * sizeof(WorkItem), hipMemcpyHostToDevice,
                        c->stream));
  // rates start at zero (a calloc'd rate vector in the reference)
  std::vector<Rate64> zero(n);
  std::memset(zero.data(), 0, n * sizeof(Rate64));
  PW_HIP(hipMemcpyAsync(c->k[0], zero.data(), n * sizeof(Rate64), hipMemcpyHostToDevice,
                        c->stream));
  PW_HIP(hipStreamSynchronize(c->stream));
  c->rates = c->k[0];
  c->zc.assign(c->nsys(), 0.0);
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess && prop.multiProcessorCount > 0)
      c->num_cus = prop.multiProcessorCount;
  }
  c->pending.clear();
  c->pending.shrink_to_fit();
  c->built = true;
  return PW_OK;
}

int kind_of(int precision, int* kind) {
  switch (precision) {
    case PW_FP64: *kind = pw::kPairFP64; return PW_OK;
    case PW_FP64_REF: *kind = pw::kPairFP64Exact; return PW_OK;
    case PW_FP32C: *kind = pw::kPairFP32C; return PW_OK;
    case PW_LEARNED: *kind = kKindLearned; return PW_OK;
    default:
      return fail(PW_ERR_ARG, "unknown precision " + std::to_string(precision));
  }
}

// Work items for a launch split into `parts` source ranges per target
// block (parts == 1: one item per block writing its rates directly).
// Split points fall on tile boundaries of the block's system, so the
// diagonal tile (and its self pairs) stays inside one part.  Items are
// ordered by cost, largest first, and among equal costs part by part: the
// items in flight at any time then read one source range, which the L2 of
// every XCD holds (a part of a 1M-particle wake split 16 ways is 3 MB of
// FP32 source records), instead of every range at once.
std::vector<WorkItem> make_items(const pw_ctx* c, int parts) {
  std::vector<WorkItem> items;
  for (const auto& s : c->systems) {
    const int64_t ntile = (s.b - s.a + kTargetsPerItem - 1) / kTargetsPerItem;
    for (int p = 0; p < parts; ++p) {
      const int64_t a = s.a + kTargetsPerItem * (ntile * p / parts);
      const int64_t b = std::min<int64_t>(s.b, s.a + kTargetsPerItem * (ntile * (p + 1) / parts));
      for (int64_t t0 = s.a; t0 < s.b; t0 += kTargetsPerItem) {
        const int32_t t1 = static_cast<int32_t>(std::min<int64_t>(t0 + kTargetsPerItem, s.b));
        items.push_back({static_cast<int32_t>(t0), t1, static_cast<int32_t>(a),
                         static_cast<int32_t>(std::max(a, b)), parts == 1 ? -1 : p, 0, chunk_of(c, a), 0});
      }
    }
  }
  std::stable_sort(items.begin(), items.end(), [](const WorkItem& x, const WorkItem& y) {
    return int64_t(x.s1 - x.s0) * (x.t1 - x.t0) > int64_t(y.s1 - y.s0) * (y.t1 - y.t0);
  });
  return items;
}

// Captured steps are replayed only while nothing they were captured with
// has changed; the stream drains first, so no replay is in flight.
void drop_graphs(pw_ctx* c) {
  if (c->graphs.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.second);
  c->graphs.clear();
}

int ensure_kind(pw_ctx* c, int kind) {
  if (kind == kKindLearned) return PW_OK;  // the tree path sets itself up
  auto& k = c->ks[kind];
  // the scalar-fed FP64 kernel's source records, allocated before any
  // step is captured (no allocation may happen inside a capture)
  if (k.cfg.scalar_feed && kind == pw::kPairFP64 && !c->d_src64 && dalloc(&c->d_src64, c->n) != hipSuccess) {
    (void)hipGetLastError();
    return fail(PW_ERR_NOMEM, "scalar-fed FP64 source records");
  }
  if (k.built_parts == k.cfg.parts && k.d_items) return PW_OK;
  std::vector<WorkItem> items = make_items(c, k.cfg.parts);
  PW_HIP(hipStreamSynchronize(c->stream));
  drop_graphs(c);  // captured steps point at the buffers replaced here
  (void)hipFree(k.d_items);
  k.d_items = nullptr;
  if (dalloc(&k.d_items, items.size()) != hipSuccess) return fail(PW_ERR_NOMEM, "work items");
  PW_HIP(hipMemcpy(k.d_items, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
  k.nitems = static_cast<int>(items.size());
  k.built_parts = k.cfg.parts;
  if (k.cfg.parts > 1 && c->partial_parts < k.cfg.parts) {
    (void)hipFree(c->partial);
    c->partial = nullptr;
    c->partial_parts = 0;
    if (dalloc(&c->partial, static_cast<size_t>(k.cfg.parts) * c->n) != hipSuccess)
      return fail(PW_ERR_NOMEM, "partial-sum buffer");
    c->partial_parts = k.cfg.parts;
  }
  return PW_OK;
}

// Blocks of wake pairs for the per-wake hooks: pairs (a, b) with a <= b,
// a == b the self pairs of wake a, a < b the cross pairs of a and b.
int check_blocks(const pw_ctx* c, const int* pairs, int npairs) {
  if (npairs < 0 || (npairs > 0 && !pairs)) return fail(PW_ERR_ARG, "null or negative wake-pair list");
  const int nw = static_cast<int>(c->wn.size());
  for (int k = 0; k < npairs; ++k) {
    const int a = pairs[2 * k], b = pairs[2 * k + 1];
    if (a < 0 || b < a || b >= nw) return fail(PW_ERR_ARG, "wake pairs need 0 <= a <= b < wakes");
    if (a != b && !c->coupled) return fail(PW_ERR_STATE, "cross-wake pairs need a coupled context");
  }
  return PW_OK;
}

// targets of wake a (256 at a time) against the sources of wake b
void append_items(const pw_ctx* c, int a, int b, int nself, std::vector<WorkItem>& out) {
  for (int64_t t0 = c->woff[a]; t0 < c->woff[a + 1]; t0 += kTargetsPerItem)
    out.push_back({static_cast<int32_t>(t0), static_cast<int32_t>(std::min<int64_t>(t0 + kTargetsPerItem, c->woff[a + 1])),
                   static_cast<int32_t>(c->woff[b]), static_cast<int32_t>(c->woff[b + 1]), -1, nself,
                   chunk_of(c, c->woff[b]), 0});
}

// the pair-kernel kind a precision launches under the context's kernel and
// summation settings
int launch_kind_of(const pw_ctx* c, int kind) {
  if (c->smoothing == PW_KERNEL_HIGHORDER) return kind == pw::kPairFP32C ? pw::kPairFP32CHO : pw::kPairFP64HO;
  if (kind == pw::kPairFP32C && !c->compensated) return pw::kPairFP32CPlain;
  return kind;
}

// One right-hand side f(state) -> out, with optional event timing.
int rhs(pw_ctx* c, int kind, const Rec64* state, Rate64* out) {
  if (kind == kKindLearned) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->profiling) {
      PW_TRY(pwi::prof_begin(c, &e0, &e1));
      PW_HIP(hipEventRecord(e0, c->stream));
    }
    PW_TRY(pwi::ff_rhs(c, state, out));
    if (c->profiling) {
      PW_HIP(hipEventRecord(e1, c->stream));
      c->ev_used.emplace_back(e0, e1);
    }
    PW_HIP(pw::launch_check_finite(out, static_cast<int>(c->n), c->d_flag, c->stream));
    return PW_OK;
  }
  if (kind == pw::kPairFP32C && !c->compensated && c->smoothing == PW_KERNEL_GAUSSIAN) {
    // plain FP32 sums: one item per 256-target block over all its sources
    pw::PairArgs args{state, c->pair_items, c->npair, static_cast<int>(c->n), out, nullptr, c->d_queue, c->nu,
                      c->d_chunks, c->d_src32, nullptr, c->nchunks};
    PW_HIP(pw::launch_pair(pw::kPairFP32CPlain, pw::PairConfig{}, args, c->npair, c->stream));
    PW_HIP(pw::launch_check_finite(out, static_cast<int>(c->n), c->d_flag, c->stream));
    return PW_OK;
  }
  PW_TRY(ensure_kind(c, kind));
  const auto& k = c->ks[kind];
  pw::PairArgs args;
  args.state = state;
  args.items = k.d_items;
  args.nitems = k.nitems;
  args.n = static_cast<int>(c->n);
  args.rates = out;
  args.partial = c->partial;
  args.queue = c->d_queue;
  args.nu = c->nu;
  args.chunks = c->d_chunks;
  args.src = c->d_src32;
  args.src64 = nullptr;
  args.nchunks = c->nchunks;
  if (k.cfg.scalar_feed && launch_kind_of(c, kind) == pw::kPairFP64) args.src64 = c->d_src64;  // ensure_kind made it
  // the high-order kernel replaces the Gaussian kinds (pw_set_kernel)
  const int launch_kind = launch_kind_of(c, kind);
  const int grid = c->num_cus * pw::pair_kernel_waves(launch_kind, k.cfg);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->profiling) {
    PW_TRY(pwi::prof_begin(c, &e0, &e1));
    PW_HIP(hipEventRecord(e0, c->stream));
  }
  PW_HIP(pw::launch_pair(launch_kind, k.cfg, args, grid, c->stream));
  if (c->profiling) {
    PW_HIP(hipEventRecord(e1, c->stream));
    c->ev_used.emplace_back(e0, e1);
  }
  PW_HIP(pw::launch_check_finite(out, static_cast<int>(c->n), c->d_flag, c->stream));
  return PW_OK;
}

int check_flag(pw_ctx* c, const char* what) {
  PW_TRY(pwi::ff_check(c));
  int flag = 0;
  PW_HIP(hipMemcpyAsync(&flag, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  PW_HIP(hipStreamSynchronize(c->stream));
  if (flag) return fail(PW_ERR_NONFINITE, std::string(what) + ": non-finite rate (coincident particles?)");
  return PW_OK;
}

int collect_profile(pw_ctx* c) {
  if (c->ev_used.empty()) return PW_OK;
  PW_HIP(hipStreamSynchronize(c->stream));
  for (auto& e : c->ev_used) {
    float ms = 0.0f;
    PW_HIP(hipEventElapsedTime(&ms, e.first, e.second));
    c->prof_ms += ms;
    c->prof_launches += 1;
    c->ev_free.push_back(e);
  }
  c->ev_used.clear();
  return PW_OK;
}

int64_t total_maxsize(const pw_ctx* c) {
  int64_t s = 0;
  for (int64_t m : c->wmax) s += 6 * m;
  return s;
}

int download_state(pw_ctx* c, std::vector<Rec64>& h) {
  h.resize(static_cast<size_t>(c->n));
  if (!c->built) {
    h = c->pending;
    return PW_OK;
  }
  PW_HIP(hipMemcpyAsync(h.data(), c->y, h.size() * sizeof(Rec64), hipMemcpyDeviceToHost,
                        c->stream));
  PW_HIP(hipStreamSynchronize(c->stream));
  return PW_OK;
}

int upload_state(pw_ctx* c, const std::vector<Rec64>& h) {
  if (!c->built) {
    c->pending = h;
    return PW_OK;
  }
  PW_HIP(hipMemcpyAsync(c->y, h.data(), h.size() * sizeof(Rec64), hipMemcpyHostToDevice,
                        c->stream));
  PW_HIP(hipStreamSynchronize(c->stream));
  return PW_OK;
}

int download_rates(pw_ctx* c, std::vector<Rate64>& h) {
  h.resize(static_cast<size_t>(c->n));
  if (!c->built) {
    std::memset(h.data(), 0, h.size() * sizeof(Rate64));
    return PW_OK;
  }
  PW_HIP(hipMemcpyAsync(h.data(), c->rates, h.size() * sizeof(Rate64), hipMemcpyDeviceToHost,
                        c->stream));
  PW_HIP(hipStreamSynchronize(c->stream));
  return PW_OK;
}

}  // namespace

namespace pwi {

int ensure_built(pw_ctx* c) { return ::ensure_built(c); }

int prof_begin(pw_ctx* c, hipEvent_t* e0, hipEvent_t* e1) {
  if (c->ev_free.empty()) {
    PW_HIP(hipEventCreate(e0));
    PW_HIP(hipEventCreate(e1));
  } else {
    *e0 = c->ev_free.back().first;
    *e1 = c->ev_free.back().second;
    c->ev_free.pop_back();
  }
  return PW_OK;
}

}  // namespace pwi

extern "C" {

const char* pw_version(void) { return "pawan-mi355x 0.1 (gfx950)"; }

const char* pw_last_error(void) { return pwi::g_error.c_str(); }

int pw_device_count(int* count) {
  if (!count) return fail(PW_ERR_ARG, "null count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    *count = 0;
    return fail(PW_ERR_HIP, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
  }
  *count = n;
  return PW_OK;
}

int pw_create(pw_ctx** out, int device) {
  if (!out) return fail(PW_ERR_ARG, "null output pointer");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(PW_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= ndev)
    return fail(PW_ERR_ARG, "device index out of range");
  pw_ctx* c = new (std::nothrow) pw_ctx();
  if (!c) return fail(PW_ERR_NOMEM, "host allocation failed");
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(PW_ERR_HIP, "stream creation failed");
  }
  const int rc = resolve_noalloc_staging(device);
  if (rc != PW_OK) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  *out = c;
  return PW_OK;
}

int pw_destroy(pw_ctx* c) {
  if (!c) return PW_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  release(c);
  delete c;
  return PW_OK;
}

int pw_add_wake(pw_ctx* c, int64_t n, int64_t maxn, const double* pos, const double* alpha,
                const double* radius, const double* volume, int* wake) {
  PW_CTX(c);
  if (c->built) return fail(PW_ERR_STATE, "wakes must be added before the first solve/step");
  if (n < 0 || maxn < n) return fail(PW_ERR_ARG, "need 0 <= n <= maxn");
  if (n > 0 && (!pos || !alpha || !radius || !volume))
    return fail(PW_ERR_ARG, "null particle array");
  if (c->n + n > INT32_MAX / 2) return fail(PW_ERR_UNSUPPORTED, "too many particles");
  for (int64_t i = 0; i < n; ++i) {
    Rec64 r;
    r.x = pos[3 * i];
    r.y = pos[3 * i + 1];
    r.z = pos[3 * i + 2];
    r.s2 = radius[i] * radius[i];  // gsl_pow_2(s), interaction_utils.h:476
    r.ax = alpha[3 * i];
    r.ay = alpha[3 * i + 1];
    r.az = alpha[3 * i + 2];
    r.vol = volume[i];
    c->pending.push_back(r);
    c->radius.push_back(radius[i]);
    c->volume.push_back(volume[i]);
  }
  c->wn.push_back(n);
  c->wmax.push_back(maxn);
  c->n += n;
  c->woff.push_back(c->n);
  if (wake) *wake = static_cast<int>(c->wn.size()) - 1;
  return PW_OK;
}

int pw_set_nu(pw_ctx* c, double nu) {
  PW_CTX(c);
  if (!std::isfinite(nu) || nu < 0) return fail(PW_ERR_ARG, "nu must be finite and >= 0");
  if (nu != c->nu) {
    drop_graphs(c);  // captured steps carry nu in their kernel arguments
    c->nu = nu;
  }
  return PW_OK;
}

int pw_get_nu(pw_ctx* c, double* nu) {
  PW_CTX(c);
  if (!nu) return fail(PW_ERR_ARG, "null output");
  *nu = c->nu;
  return PW_OK;
}

int pw_set_coupled(pw_ctx* c, int coupled) {
  PW_CTX(c);
  if (c->built) return fail(PW_ERR_STATE, "coupling must be set before the first solve/step");
  c->coupled = coupled != 0;
  return PW_OK;
}

int pw_set_kernel(pw_ctx* c, int kernel) {
  PW_CTX(c);
  if (kernel != PW_KERNEL_GAUSSIAN && kernel != PW_KERNEL_HIGHORDER)
    return fail(PW_ERR_ARG, "kernel must be PW_KERNEL_GAUSSIAN or PW_KERNEL_HIGHORDER");
  if (kernel != c->smoothing) {
    drop_graphs(c);  // captured steps hold the other kernel's launches
    c->smoothing = kernel;
  }
  return PW_OK;
}

int pw_set_compensation(pw_ctx* c, int enable) {
  PW_CTX(c);
  if ((enable != 0) != c->compensated) {
    drop_graphs(c);  // captured steps hold the other summation's launches
    c->compensated = enable != 0;
  }
  return PW_OK;
}

int pw_get_kernel(pw_ctx* c, int* kernel) {
  PW_CTX(c);
  if (!kernel) return fail(PW_ERR_ARG, "null output");
  *kernel = c->smoothing;
  return PW_OK;
}

int pw_get_size(pw_ctx* c, int64_t* num_particles, int* num_wakes) {
  PW_CTX(c);
  if (num_particles) *num_particles = c->n;
  if (num_wakes) *num_wakes = static_cast<int>(c->wn.size());
  return PW_OK;
}

int pw_set_wake_state(pw_ctx* c, int w, const double* pos, const double* alpha) {
  PW_CTX(c);
  if (w < 0 || w >= static_cast<int>(c->wn.size())) return fail(PW_ERR_ARG, "wake index out of range");
  if (c->wn[w] > 0 && (!pos || !alpha)) return fail(PW_ERR_ARG, "null array");
  std::vector<Rec64> h;
  PW_TRY(download_state(c, h));
  for (int64_t i = 0; i < c->wn[w]; ++i) {
    Rec64& r = h[c->woff[w] + i];
    r.x = pos[3 * i]; r.y = pos[3 * i + 1]; r.z = pos[3 * i + 2];
    r.ax = alpha[3 * i]; r.ay = alpha[3 * i + 1]; r.az = alpha[3 * i + 2];
  }
  return upload_state(c, h);
}

int pw_set_wake_volume(pw_ctx* c, int w, const double* volume) {
  PW_CTX(c);
  if (w < 0 || w >= static_cast<int>(c->wn.size())) return fail(PW_ERR_ARG, "wake index out of range");
  if (c->wn[w] > 0 && !volume) return fail(PW_ERR_ARG, "null array");
  std::vector<Rec64> h;
  PW_TRY(download_state(c, h));
  for (int64_t i = 0; i < c->wn[w]; ++i) {
    h[c->woff[w] + i].vol = volume[i];
    c->volume[c->woff[w] + i] = volume[i];
  }
  return upload_state(c, h);
}

int pw_get_wake_state(pw_ctx* c, int w, double* pos, double* alpha) {
  PW_CTX(c);
  if (w < 0 || w >= static_cast<int>(c->wn.size())) return fail(PW_ERR_ARG, "wake index out of range");
  std::vector<Rec64> h;
  PW_TRY(download_state(c, h));
  for (int64_t i = 0; i < c->wn[w]; ++i) {
    const Rec64& r = h[c->woff[w] + i];
    if (pos) { pos[3 * i] = r.x; pos[3 * i + 1] = r.y; pos[3 * i + 2] = r.z; }
    if (alpha) { alpha[3 * i] = r.ax; alpha[3 * i + 1] = r.ay; alpha[3 * i + 2] = r.az; }
  }
  return PW_OK;
}

int pw_get_wake_rates(pw_ctx* c, int w, double* vel, double* retvor) {
  PW_CTX(c);
  if (w < 0 || w >= static_cast<int>(c->wn.size())) return fail(PW_ERR_ARG, "wake index out of range");
  std::vector<Rate64> h;
  PW_TRY(download_rates(c, h));
  for (int64_t i = 0; i < c->wn[w]; ++i) {
    const Rate64& r = h[c->woff[w] + i];
    if (vel) { vel[3 * i] = r.ux; vel[3 * i + 1] = r.uy; vel[3 * i + 2] = r.uz; }
    if (retvor) { retvor[3 * i] = r.wx; retvor[3 * i + 1] = r.wy; retvor[3 * i + 2] = r.wz; }
  }
  return PW_OK;
}

int pw_set_states(pw_ctx* c, const double* state, int64_t len) {
  PW_CTX(c);
  if (!state) return fail(PW_ERR_ARG, "null state");
  if (len != total_maxsize(c)) return fail(PW_ERR_ARG, "state length != sum of 6*maxn");
  std::vector<Rec64> h;
  PW_TRY(download_state(c, h));
  int64_t base = 0;
  for (size_t w = 0; w < c->wn.size(); ++w) {
    const double* p = state + base;                 // pos block, maxn x 3
    const double* a = state + base + 3 * c->wmax[w];  // alpha block (wake.cpp:860-866)
    for (int64_t i = 0; i < c->wn[w]; ++i) {
      Rec64& r = h[c->woff[w] + i];
      r.x = p[3 * i]; r.y = p[3 * i + 1]; r.z = p[3 * i + 2];
      r.ax = a[3 * i]; r.ay = a[3 * i + 1]; r.az = a[3 * i + 2];
    }
    base += 6 * c->wmax[w];
  }
  return upload_state(c, h);
}

int pw_get_states(pw_ctx* c, double* state, int64_t len) {
  PW_CTX(c);
  if (!state) return fail(PW_ERR_ARG, "null state");
  if (len != total_maxsize(c)) return fail(PW_ERR_ARG, "state length != sum of 6*maxn");
  std::vector<Rec64> h;
  PW_TRY(download_state(c, h));
  std::memset(state, 0, sizeof(double) * static_cast<size_t>(len));
  int64_t base = 0;
  for (size_t w = 0; w < c->wn.size(); ++w) {
    double* p = state + base;
    double* a = state + base + 3 * c->wmax[w];
    for (int64_t i = 0; i < c->wn[w]; ++i) {
      const Rec64& r = h[c->woff[w] + i];
      p[3 * i] = r.x; p[3 * i + 1] = r.y; p[3 * i + 2] = r.z;
      a[3 * i] = r.ax; a[3 * i + 1] = r.ay; a[3 * i + 2] = r.az;
    }
    base += 6 * c->wmax[w];
  }
  return PW_OK;
}

int pw_get_rates(pw_ctx* c, double* rate, int64_t len) {
  PW_CTX(c);
  if (!rate) return fail(PW_ERR_ARG, "null rate");
  if (len != total_maxsize(c)) return fail(PW_ERR_ARG, "rate length != sum of 6*maxn");
  std::vector<Rate64> h;
  PW_TRY(download_rates(c, h));
  std::memset(rate, 0, sizeof(double) * static_cast<size_t>(len));
  int64_t base = 0;
  for (size_t w = 0; w < c->wn.size(); ++w) {
    double* u = rate + base;
    double* q = rate + base + 3 * c->wmax[w];  // wake.cpp:876
    for (int64_t i = 0; i < c->wn[w]; ++i) {
      const Rate64& r = h[c->woff[w] + i];
      u[3 * i] = r.ux; u[3 * i + 1] = r.uy; u[3 * i + 2] = r.uz;
      q[3 * i] = r.wx; q[3 * i + 1] = r.wy; q[3 * i + 2] = r.wz;
    }
    base += 6 * c->wmax[w];
  }
  return PW_OK;
}

int pw_velocity_at(pw_ctx* c, const double* points, int64_t m, double* u) {
  PW_CTX(c);
  if (m < 0 || m > (int64_t(1) << 28)) return fail(PW_ERR_ARG, "m must be in [0, 2^28]");
  if (m == 0) return PW_OK;
  if (!points || !u) return fail(PW_ERR_ARG, "null points or output");
  PW_TRY(ensure_built(c));
  double *d_pts = nullptr, *d_u = nullptr;
  const size_t len = 3 * static_cast<size_t>(m);
  if (dalloc(&d_pts, len) != hipSuccess || dalloc(&d_u, len) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(d_pts);
    return fail(PW_ERR_NOMEM, "point buffers");
  }
  hipError_t e = hipMemcpyAsync(d_pts, points, len * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = pw::launch_point_velocity(c->y, static_cast<int>(c->n), d_pts, static_cast<int>(m), d_u,
                                  c->smoothing == PW_KERNEL_HIGHORDER, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(u, d_u, len * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  else (void)hipStreamSynchronize(c->stream);  // nothing queued may still use the buffers
  (void)hipFree(d_pts);
  (void)hipFree(d_u);
  if (e != hipSuccess) return fail(PW_ERR_HIP, std::string("pw_velocity_at: ") + hipGetErrorString(e));
  return PW_OK;
}

// pw_velocity_at without its per-call buffers: the points and velocities
// pass through the device's static staging array.  Calls on one GPU share
// that array, so calls that overlap on one GPU (several host threads) race.
int pw_velocity_at_noalloc(pw_ctx* c, const double* points, int64_t m, double* u) {
  PW_CTX(c);
  if (m < 0 || m > kNoAllocPoints)
    return fail(PW_ERR_ARG, "m must be in [0, " + std::to_string(kNoAllocPoints) +
                                "], the capacity of the static staging buffer");
  if (m == 0) return PW_OK;
  if (!points || !u) return fail(PW_ERR_ARG, "null points or output");
  if (!c->built)
    return fail(PW_ERR_STATE, "the context's particles are not on the device yet: its first pw_solve, pw_step "
                              "or pw_velocity_at puts them there (and allocates)");
  double* st = noalloc_staging(c->device);
  if (!st) return fail(PW_ERR_STATE, "staging buffer of this device not resolved");
  double *d_pts = st, *d_u = st + 3 * kNoAllocPoints;
  const size_t bytes = 3 * static_cast<size_t>(m) * sizeof(double);
  hipError_t e = hipMemcpyAsync(d_pts, points, bytes, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = pw::launch_point_velocity(c->y, static_cast<int>(c->n), d_pts, static_cast<int>(m), d_u,
                                  c->smoothing == PW_KERNEL_HIGHORDER, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(u, d_u, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  else (void)hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(PW_ERR_HIP, std::string("pw_velocity_at_noalloc: ") + hipGetErrorString(e));
  return PW_OK;
}

int pw_vring_particles(double R, double core, int nlayers, int nazi, double sig, double* pos,
                       double* radius, double* volume) {
  (void)core;  // sets only the ideal profile (pw_vring_resolve)
  if (nlayers < 0 || nazi < 1 || (2 * nlayers + 1) * (2 * nlayers + 1) > pw::kVringMaxCells)
    return fail(PW_ERR_ARG, "nlayers must be in [0, 15] and nazi >= 1");
  if (!(std::isfinite(R) && R > 0.0) || !(std::isfinite(sig) && sig > 0.0))
    return fail(PW_ERR_ARG, "ring radius and smoothing radius must be finite and positive");
  if (!pos || !radius || !volume) return fail(PW_ERR_ARG, "null output");
  // the particle layout of __vring (vring.cpp:50-100): one centre particle
  // and layers of 8j particles per station, the cell volumes of its
  // annular sectors
  const double r0 = 0.35 / (2 * nlayers + 1);
  const double dpsi = 2.0 * M_PI / nazi;
  int64_t q = 0;
  for (int a = 0; a < nazi; ++a) {
    const double psi = a * dpsi, sp = std::sin(psi), cp = std::cos(psi);
    pos[3 * q] = R * cp;
    pos[3 * q + 1] = R * sp;
    pos[3 * q + 2] = 0.0;
    radius[q] = sig;
    volume[q] = M_PI * (r0 * r0) * R * dpsi;
    ++q;
    for (int j = 1; j <= nlayers; ++j) {
      const double jj = static_cast<double>(j) * j;
      const double r = r0 * (12. * jj + 1.0) / (6. * j);
      const int ncell = 8 * j;
      const double dth = 2. * M_PI / ncell;
      for (int k = 0; k < ncell; ++k) {
        const double th = k * dth;
        pos[3 * q] = (R + r * std::cos(th)) * cp;
        pos[3 * q + 1] = (R + r * std::cos(th)) * sp;
        pos[3 * q + 2] = r * std::sin(th);
        radius[q] = sig;
        volume[q] = 2.0 * r0 * dpsi * (2.0 * dth * j * R * r0 + (12.0 * jj + 1.0) * (r0 * r0) *
                                                                    (std::sin(th + dth) - std::sin(th)) / 3.0);
        ++q;
      }
    }
  }
  return PW_OK;
}

int pw_vring_resolve(pw_ctx* c, int wake, int nlayers, int nazi, double R, double core, double* strength,
                     double* influence) {
  PW_CTX(c);
  if (wake < 0 || wake >= static_cast<int>(c->wn.size())) return fail(PW_ERR_ARG, "wake index out of range");
  if (nlayers < 0 || nazi < 1 || (2 * nlayers + 1) * (2 * nlayers + 1) > pw::kVringMaxCells)
    return fail(PW_ERR_ARG, "nlayers must be in [0, 15] and nazi >= 1");
  if (!(std::isfinite(core) && core > 0.0) || !(std::isfinite(R) && R > 0.0))
    return fail(PW_ERR_ARG, "radius and core must be finite and positive");
  const int m = (2 * nlayers + 1) * (2 * nlayers + 1);
  if (c->wn[wake] != static_cast<int64_t>(m) * nazi)
    return fail(PW_ERR_ARG, "the wake does not hold nazi (2 nlayers + 1)^2 particles");
  // a zero or non-finite smoothing radius makes every influence entry 0/0
  for (int64_t q = c->woff[wake]; q < c->woff[wake + 1]; ++q)
    if (!(std::isfinite(c->radius[q]) && c->radius[q] > 0.0))
      return fail(PW_ERR_ARG, "the ring's smoothing radii must be finite and positive");
  PW_TRY(ensure_built(c));
  // __vring::getIdealRates (vring.cpp:223-242): the Gaussian core profile
  std::vector<double> ideal(m);
  {
    const double mag0 = 0.5 / (M_PI * core * core), r0 = 0.35 / (2 * nlayers + 1);
    int idx = 0;
    ideal[idx++] = mag0;
    for (int j = 1; j <= nlayers; ++j) {
      const double r = r0 * (12. * (static_cast<double>(j) * j) + 1.0) / (6. * j);
      const int ncell = 8 * j;
      const double dth = 2. * M_PI / ncell, decay = std::exp(-0.5 * r * r / (core * core));
      for (int k = 0; k < ncell; ++k) ideal[idx++] = mag0 * (1.0 + r * std::cos(k * dth) / R) * decay;
    }
  }
  double *d_M = nullptr, *d_lu = nullptr, *d_b = nullptr, *d_s = nullptr;
  int* d_flag = nullptr;
  const size_t mm = static_cast<size_t>(m) * m;
  if (dalloc(&d_M, mm) != hipSuccess || dalloc(&d_lu, mm) != hipSuccess || dalloc(&d_b, m) != hipSuccess ||
      dalloc(&d_s, m) != hipSuccess || dalloc(&d_flag, 1 + static_cast<size_t>(m)) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(d_M); (void)hipFree(d_lu); (void)hipFree(d_b); (void)hipFree(d_s); (void)hipFree(d_flag);
    return fail(PW_ERR_NOMEM, "resolve buffers");
  }
  int flag = 0;
  hipError_t e = hipMemcpyAsync(d_b, ideal.data(), m * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = pw::launch_vring_resolve(c->y, static_cast<int>(c->woff[wake]), m, nazi,
                                 c->smoothing == PW_KERNEL_HIGHORDER, d_M, d_lu, d_b, d_s, d_flag, c->stream);
  if (e == hipSuccess && strength)
    e = hipMemcpyAsync(strength, d_s, m * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && influence)
    e = hipMemcpyAsync(influence, d_M, mm * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&flag, d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  else (void)hipStreamSynchronize(c->stream);  // nothing queued may still use the buffers
  (void)hipFree(d_M); (void)hipFree(d_lu); (void)hipFree(d_b); (void)hipFree(d_s); (void)hipFree(d_flag);
  if (e != hipSuccess) return fail(PW_ERR_HIP, std::string("pw_vring_resolve: ") + hipGetErrorString(e));
  if (flag) return fail(PW_ERR_NONFINITE, "pw_vring_resolve: singular influence matrix");
  return PW_OK;
}

int pw_solve(pw_ctx* c, int precision) {
  PW_CTX(c);
  int kind = 0;
  PW_TRY(kind_of(precision, &kind));
  PW_TRY(ensure_built(c));
  PW_HIP(pw::launch_fill_int(c->d_flag, 1, 0, c->stream));
  PW_TRY(rhs(c, kind, c->y, c->k[0]));
  c->rates = c->k[0];
  return check_flag(c, "pw_solve");
}

}  // extern "C"

namespace {

// The kernels of one step of `scheme` on the context stream.
int enqueue_step(pw_ctx* c, int scheme, int kind, double dt) {
  const int n = static_cast<int>(c->n);
  hipStream_t st = c->stream;
  Rate64** k = c->k;
  if (scheme == PW_RK4) {
    // rk4.cpp:20-66
    PW_TRY(rhs(c, kind, c->y, k[0]));
    PW_HIP(pw::launch_stage(c->y, k[0], 0.5 * dt, c->ys, n, st));
    PW_TRY(rhs(c, kind, c->ys, k[1]));
    PW_HIP(pw::launch_stage(c->y, k[1], 0.5 * dt, c->ys, n, st));
    PW_TRY(rhs(c, kind, c->ys, k[2]));
    PW_HIP(pw::launch_stage(c->y, k[2], dt, c->ys, n, st));
    PW_TRY(rhs(c, kind, c->ys, k[3]));
    PW_HIP(pw::launch_rk4_final(c->y, k[0], k[1], k[2], k[3], dt / 6., dt / 3., n, st));
  } else if (scheme == PW_RK3) {
    // rk3.cpp:20-55
    PW_TRY(rhs(c, kind, c->y, k[0]));
    PW_HIP(pw::launch_stage(c->y, k[0], 0.5 * dt, c->ys, n, st));
    PW_TRY(rhs(c, kind, c->ys, k[1]));
    PW_HIP(pw::launch_rk3_mid(c->y, k[0], k[1], dt, 2. * dt, c->ys, n, st));
    PW_TRY(rhs(c, kind, c->ys, k[2]));
    PW_HIP(pw::launch_rk3_final(c->y, k[0], k[1], k[2], dt / 6.0, dt * 2.0 / 3.0, n, st));
  } else {
    // integration.cpp:126-136
    PW_TRY(rhs(c, kind, c->y, k[0]));
    PW_HIP(pw::launch_euler(c->y, k[0], dt, n, st));
  }
  return PW_OK;
}

Rate64* last_rates(pw_ctx* c, int scheme) {
  return scheme == PW_RK4 ? c->k[3] : scheme == PW_RK3 ? c->k[2] : c->k[0];
}

// One step captured as a HIP graph (per scheme, kind and dt), replayed for
// every step: one launch instead of up to 13.
int step_graph(pw_ctx* c, int scheme, int kind, double dt, hipGraphExec_t* out) {
  uint64_t bits;
  std::memcpy(&bits, &dt, sizeof bits);
  const auto key = std::make_tuple(scheme, kind, bits);
  auto it = c->graphs.find(key);
  if (it != c->graphs.end()) {
    *out = it->second;
    return PW_OK;
  }
  PW_TRY(ensure_kind(c, kind));  // allocations happen outside the capture
  hipGraph_t g = nullptr;
  PW_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_step(c, scheme, kind, dt);
  hipGraph_t captured = nullptr;
  const hipError_t ec = hipStreamEndCapture(c->stream, &captured);
  if (rc != PW_OK || ec != hipSuccess) {
    // a failed capture leaves its error as the thread's last error: clear
    // it, so that the next launch's check does not report it again
    (void)hipGetLastError();
    if (captured) (void)hipGraphDestroy(captured);
    if (rc != PW_OK) return rc;
    return fail(PW_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
  }
  g = captured;
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) return fail(PW_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  c->graphs[key] = exec;
  *out = exec;
  return PW_OK;
}

}  // namespace

extern "C" {

int pw_step(pw_ctx* c, int scheme, double dt, int precision, int64_t nsteps) {
  PW_CTX(c);
  int kind = 0;
  PW_TRY(kind_of(precision, &kind));
  if (scheme != PW_EULER && scheme != PW_RK3 && scheme != PW_RK4)
    return fail(PW_ERR_ARG, "scheme must be PW_EULER, PW_RK3 or PW_RK4");
  if (!std::isfinite(dt)) return fail(PW_ERR_ARG, "dt must be finite");
  if (nsteps < 0) return fail(PW_ERR_ARG, "nsteps must be >= 0");
  PW_TRY(ensure_built(c));
  PW_TRY(ensure_kind(c, kind));
  PW_HIP(pw::launch_fill_int(c->d_flag, 1, 0, c->stream));
  // the tree path re-sorts the particles every stage (host-side sort
  // set-up), so its steps are launched eagerly
  if (c->use_graph && !c->profiling && nsteps > 0 && kind != kKindLearned) {
    hipGraphExec_t exec = nullptr;
    PW_TRY(step_graph(c, scheme, kind, dt, &exec));
    for (int64_t it = 0; it < nsteps; ++it) PW_HIP(hipGraphLaunch(exec, c->stream));
  } else {
    for (int64_t it = 0; it < nsteps; ++it) PW_TRY(enqueue_step(c, scheme, kind, dt));
  }
  if (nsteps > 0) c->rates = last_rates(c, scheme);
  return check_flag(c, "pw_step");
}

int pw_interact_wakes(pw_ctx* c, int precision, const int* pairs, int npairs) {
  PW_CTX(c);
  int kind = 0;
  PW_TRY(kind_of(precision, &kind));
  if (kind == kKindLearned) return fail(PW_ERR_UNSUPPORTED, "the tree path evaluates the whole system at once");
  PW_TRY(check_blocks(c, pairs, npairs));
  PW_TRY(ensure_built(c));
  std::vector<WorkItem> items;
  std::vector<int> first(npairs + 1, 0);
  for (int k = 0; k < npairs; ++k) {
    const int a = pairs[2 * k], b = pairs[2 * k + 1];
    append_items(c, a, b, 0, items);
    if (a != b) append_items(c, b, a, 0, items);
    first[k + 1] = static_cast<int>(items.size());
  }
  const int n = static_cast<int>(c->n);
  Rate64* acc = c->k[0];
  Rate64* blk = c->k[1];
  PW_HIP(pw::launch_fill_int(c->d_flag, 1, 0, c->stream));
  PW_HIP(pw::launch_zero_doubles(reinterpret_cast<double*>(acc), 6 * static_cast<int64_t>(n), c->stream));
  WorkItem* d_items = nullptr;
  if (!items.empty() && dalloc(&d_items, items.size()) != hipSuccess) {
    (void)hipGetLastError();
    return fail(PW_ERR_NOMEM, "work items");
  }
  hipError_t e = items.empty() ? hipSuccess
                               : hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(WorkItem),
                                                hipMemcpyHostToDevice, c->stream);
  const int lk = launch_kind_of(c, kind);
  for (int k = 0; k < npairs && e == hipSuccess; ++k) {
    // one block at a time: its items have distinct targets and write their
    // rates into blk, which is then added to the targets' totals
    pw::PairArgs args{c->y,     d_items + first[k], first[k + 1] - first[k], n, blk, nullptr, c->d_queue, c->nu,
                      c->d_chunks, c->d_src32,        nullptr, c->nchunks};
    e = pw::launch_pair(lk, pw::PairConfig{}, args, args.nitems, c->stream);
    const int a = pairs[2 * k], b = pairs[2 * k + 1];
    if (e == hipSuccess)
      e = pw::launch_add_rates(acc, blk, static_cast<int>(c->woff[a]), static_cast<int>(c->woff[a + 1]), c->stream);
    if (e == hipSuccess && a != b)
      e = pw::launch_add_rates(acc, blk, static_cast<int>(c->woff[b]), static_cast<int>(c->woff[b + 1]), c->stream);
  }
  if (e == hipSuccess) e = pw::launch_check_finite(acc, n, c->d_flag, c->stream);
  (void)hipStreamSynchronize(c->stream);  // the items are freed below
  (void)hipFree(d_items);
  if (e != hipSuccess) return fail(PW_ERR_HIP, std::string("pw_interact_wakes: ") + hipGetErrorString(e));
  c->rates = acc;
  return check_flag(c, "pw_interact_wakes");
}

int pw_vorticity_field(pw_ctx* c, const int* pairs, int npairs, double* field, int64_t len) {
  PW_CTX(c);
  if (!field) return fail(PW_ERR_ARG, "null output");
  if (pairs) PW_TRY(check_blocks(c, pairs, npairs));
  PW_TRY(ensure_built(c));
  const int n = static_cast<int>(c->n);
  if (len != 3 * static_cast<int64_t>(n)) return fail(PW_ERR_ARG, "field needs 3 doubles per particle");
  // launch groups whose items have distinct targets: the whole system at
  // once, or one block of wake pairs at a time
  std::vector<WorkItem> items;
  std::vector<int> first{0};
  if (!pairs) {
    items = c->h_pair_items;
    first.push_back(static_cast<int>(items.size()));
  } else {
    for (int k = 0; k < npairs; ++k) {
      const int a = pairs[2 * k], b = pairs[2 * k + 1];
      append_items(c, a, b, 0, items);
      if (a != b) append_items(c, b, a, 0, items);
      first.push_back(static_cast<int>(items.size()));
    }
  }
  double* d_field = reinterpret_cast<double*>(c->k[1]);  // 6 doubles per particle
  WorkItem* d_items = nullptr;
  if (!items.empty() && dalloc(&d_items, items.size()) != hipSuccess) {
    (void)hipGetLastError();
    return fail(PW_ERR_NOMEM, "work items");
  }
  hipError_t e = pw::launch_zero_doubles(d_field, 3 * static_cast<int64_t>(n), c->stream);
  if (e == hipSuccess && !items.empty())
    e = hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice, c->stream);
  const bool ho = c->smoothing == PW_KERNEL_HIGHORDER;
  for (size_t g = 0; g + 1 < first.size() && e == hipSuccess; ++g)
    e = pw::launch_vorticity_field(c->y, d_items + first[g], first[g + 1] - first[g], ho, d_field, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(field, d_field, sizeof(double) * 3 * static_cast<size_t>(n), hipMemcpyDeviceToHost,
                       c->stream);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(d_items);
  if (e != hipSuccess) return fail(PW_ERR_HIP, std::string("pw_vorticity_field: ") + hipGetErrorString(e));
  return PW_OK;
}

int pw_diagnose_wakes(pw_ctx* c, int wa, int wb, double* out, int64_t out_len) {
  PW_CTX(c);
  if (!out || out_len < 16) return fail(PW_ERR_ARG, "output needs 16 doubles");
  const int pair[2] = {wa, wb};
  PW_TRY(check_blocks(c, pair, 1));
  PW_TRY(ensure_built(c));
  std::vector<WorkItem> items;
  append_items(c, wa, wb, wa == wb ? 0 : 1, items);
  constexpr int Q = pw::kDiagQuad;
  const int ni = static_cast<int>(items.size());
  // the wake's linear-sum items are contiguous in the context's list
  int l0 = 0, nl = 0;
  for (int b = 0; b < c->nlin; ++b)
    if (c->lin_wake[b] == wa) {
      if (nl == 0) l0 = b;
      ++nl;
    }
  WorkItem* d_items = nullptr;
  double* d_q = nullptr;
  if (ni > 0 && (dalloc(&d_items, items.size()) != hipSuccess || dalloc(&d_q, Q * items.size()) != hipSuccess)) {
    (void)hipGetLastError();
    (void)hipFree(d_items);
    return fail(PW_ERR_NOMEM, "diagnostic buffers");
  }
  std::vector<double> quad(Q * static_cast<size_t>(ni)), lin(11 * static_cast<size_t>(nl));
  hipError_t e = hipSuccess;
  if (ni > 0) {
    e = hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
      e = pw::launch_diag_pairs(c->y, c->d_radius, d_items, ni, d_q, c->smoothing == PW_KERNEL_HIGHORDER,
                                c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(quad.data(), d_q, quad.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  }
  if (e == hipSuccess && wa == wb && nl > 0) {
    e = pw::launch_diag_linear(c->y, c->lin_items + l0, nl, c->d_lin + 11 * static_cast<size_t>(l0), c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(lin.data(), c->d_lin + 11 * static_cast<size_t>(l0), lin.size() * sizeof(double),
                         hipMemcpyDeviceToHost, c->stream);
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(d_items);
  (void)hipFree(d_q);
  if (e != hipSuccess) return fail(PW_ERR_HIP, std::string("pw_diagnose_wakes: ") + hipGetErrorString(e));
  for (int m = 0; m < 16; ++m) out[m] = 0.0;
  if (wa == wb) {
    double v[11] = {0};
    for (int b = 0; b < nl; ++b)
      for (int m = 0; m < 11; ++m) v[m] += lin[11 * b + m];
    for (int m = 0; m < 3; ++m) {
      out[m] = v[m];                       // calculateTotalVorticity
      out[3 + m] = v[3 + m] * 0.5;         // calculateLinearImpulse
      out[6 + m] = v[6 + m] * (1.0 / 3.0); // calculateAngularImpulse
    }
    out[14] = v[9];
    out[15] = v[10];
  }
  double q[Q] = {0};
  for (int b = 0; b < ni; ++b)
    for (int m = 0; m < Q; ++m) q[m] += quad[Q * b + m];
  // ordered sum over the block's (target, source) pairs: within a wake it is
  // 2 sum_{i<j} plus the self terms; across wakes one direction, doubled as
  // in calculate*(W1, W2) (interaction.cpp:551-718)
  const double f = wa == wb ? 1.0 : 2.0;
  out[9] = f * q[0];   // helicity
  out[10] = f * q[3];  // enstrophy
  out[11] = f * q[4];  // div-free enstrophy
  out[12] = f * q[1];  // kinetic energy
  out[13] = f * q[2];  // div-free kinetic energy
  return PW_OK;
}

int pw_diagnose(pw_ctx* c, double* out, int64_t out_len) {
  PW_CTX(c);
  if (!out) return fail(PW_ERR_ARG, "null output");
  PW_TRY(ensure_built(c));
  const int ns = c->nsys();
  if (out_len < 15 * static_cast<int64_t>(ns))
    return fail(PW_ERR_ARG, "output needs 15 doubles per system");
  PW_HIP(pw::launch_diag_linear(c->y, c->lin_items, c->nlin, c->d_lin, c->stream));
  const bool ho = c->smoothing == PW_KERNEL_HIGHORDER;
  constexpr int Q = pw::kDiagQuad;
  PW_HIP(pw::launch_diag_pairs(c->y, c->d_radius, c->pair_items, c->npair, c->d_quad, ho, c->stream));
  std::vector<double> lin(11 * static_cast<size_t>(c->nlin)), quad(Q * static_cast<size_t>(c->npair));
  PW_HIP(hipMemcpyAsync(lin.data(), c->d_lin, lin.size() * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  PW_HIP(hipMemcpyAsync(quad.data(), c->d_quad, quad.size() * sizeof(double),
                        hipMemcpyDeviceToHost, c->stream));
  PW_HIP(hipStreamSynchronize(c->stream));
  const int nw = static_cast<int>(c->wn.size());
  // per-wake linear sums in block order
  std::vector<double> wl(11 * static_cast<size_t>(nw), 0.0);
  for (int b = 0; b < c->nlin; ++b)
    for (int m = 0; m < 11; ++m) wl[11 * c->lin_wake[b] + m] += lin[11 * b + m];
  // per-system quadratic sums in item order (items of one system are
  // contiguous and in target order)
  std::vector<double> sq(Q * static_cast<size_t>(ns), 0.0);
  for (int b = 0; b < c->npair; ++b)
    for (int m = 0; m < Q; ++m) sq[Q * c->pair_sys[b] + m] += quad[Q * b + m];
  for (int s = 0; s < ns; ++s) {
    double* o = out + 15 * s;
    for (int m = 0; m < 9; ++m) o[m] = 0.0;
    const int w0 = c->coupled ? 0 : s, w1 = c->coupled ? nw : s + 1;
    for (int w = w0; w < w1; ++w) {
      const double* v = &wl[11 * w];
      // calculateTotalVorticity / LinearImpulse (* 0.5) / AngularImpulse
      // (* 1/3), interaction.cpp:489-527, summed over wakes (:74-84)
      for (int m = 0; m < 3; ++m) {
        o[m] += v[m];
        o[3 + m] += v[3 + m] * 0.5;
        o[6 + m] += v[6 + m] * (1.0 / 3.0);
      }
      if (v[10] != 0.0) c->zc[s] = v[9] / v[10];  // interaction.cpp:153-155
    }
    o[9] = sq[Q * s + 0];   // helicity
    o[10] = sq[Q * s + 3];  // enstrophy (ENST == 0 for the Gaussian kernel)
    o[11] = sq[Q * s + 4];  // div-free enstrophy (ENSTF == 0 likewise)
    o[12] = sq[Q * s + 1];  // kinetic energy
    o[13] = sq[Q * s + 2];  // div-free kinetic energy
    o[14] = c->zc[s];
  }
  return PW_OK;
}

int pw_synchronize(pw_ctx* c) {
  PW_CTX(c);
  PW_HIP(hipStreamSynchronize(c->stream));
  return PW_OK;
}

int pw_set_profiling(pw_ctx* c, int enable) {
  PW_CTX(c);
  PW_TRY(collect_profile(c));
  PW_TRY(pwi::ff_collect_stages(c));
  c->profiling = enable != 0;
  c->prof_ms = 0.0;
  c->prof_launches = 0;
  pwi::ff_reset_stages(c);
  return PW_OK;
}

int pw_get_pair_timing(pw_ctx* c, double* total_ms, int64_t* launches) {
  PW_CTX(c);
  PW_TRY(collect_profile(c));
  if (total_ms) *total_ms = c->prof_ms;
  if (launches) *launches = c->prof_launches;
  return PW_OK;
}

int pw_set_pair_config(pw_ctx* c, int precision, int parts, int persistent, int variant) {
  PW_CTX(c);
  int kind = 0;
  PW_TRY(kind_of(precision, &kind));
  if (kind == kKindLearned) return fail(PW_ERR_UNSUPPORTED, "launch configuration applies to the direct-sum kinds");
  if (parts < 1 || parts > 256) return fail(PW_ERR_ARG, "parts must be in [1, 256]");
  auto& k = c->ks[kind];
  if (variant < 0 || variant > (PW_VARIANT_WAVES8 | PW_VARIANT_SCALAR_FEED | PW_VARIANT_TWO_PER_LANE))
    return fail(PW_ERR_ARG, "variant must be a combination of the PW_VARIANT_* bits");
  k.cfg.parts = parts;
  k.cfg.persistent = persistent != 0;
  k.cfg.scalar_feed = (variant & PW_VARIANT_SCALAR_FEED) != 0 && (kind == pw::kPairFP32C || kind == pw::kPairFP64);
  k.cfg.two_per_lane = (variant & PW_VARIANT_TWO_PER_LANE) != 0 && k.cfg.scalar_feed && kind == pw::kPairFP32C;
  k.cfg.waves8 = (variant & PW_VARIANT_WAVES8) != 0 && kind == pw::kPairFP32C && !k.cfg.two_per_lane;
  k.tuned = true;
  drop_graphs(c);
  return PW_OK;
}

int pw_get_pair_config(pw_ctx* c, int precision, int* parts, int* persistent, int* variant) {
  PW_CTX(c);
  int kind = 0;
  PW_TRY(kind_of(precision, &kind));
  if (kind == kKindLearned) return fail(PW_ERR_UNSUPPORTED, "launch configuration applies to the direct-sum kinds");
  const auto& k = c->ks[kind];
  if (parts) *parts = k.cfg.parts;
  if (persistent) *persistent = k.cfg.persistent ? 1 : 0;
  if (variant)
    *variant = (k.cfg.waves8 ? PW_VARIANT_WAVES8 : 0) | (k.cfg.scalar_feed ? PW_VARIANT_SCALAR_FEED : 0) |
               (k.cfg.two_per_lane ? PW_VARIANT_TWO_PER_LANE : 0);
  return PW_OK;
}

int pw_set_graph(pw_ctx* c, int enable) {
  PW_CTX(c);
  c->use_graph = enable != 0;
  if (!c->use_graph) drop_graphs(c);
  return PW_OK;
}

}  // extern "C"

namespace {

// Tuning results shared by every context of the process, keyed by the
// problem's shape; optionally also appended to / read from the text file
// named by $PAWAN_TUNE_CACHE ("kind n nsys largest parts persistent variant",
// kind the pair-kernel family: 0-2 for the Gaussian precisions, 4 and 5
// for the high-order FP64 and FP32C kernels; variant the PW_VARIANT_* bits).
using TuneKey = std::tuple<int, int64_t, int, int64_t>;
std::map<TuneKey, pw::PairConfig>& tune_cache() {
  static std::map<TuneKey, pw::PairConfig> m;
  return m;
}
// guards the cache and its file (contexts may be driven from several threads)
std::mutex& tune_mutex() {
  static std::mutex m;
  return m;
}

void load_tune_file() {
  static bool loaded = false;
  if (loaded) return;
  loaded = true;
  const char* path = std::getenv("PAWAN_TUNE_CACHE");
  if (!path) return;
  FILE* f = std::fopen(path, "r");
  if (!f) return;
  int kind, nsys, parts, pers, w8;
  long long n, largest;
  while (std::fscanf(f, "%d %lld %d %lld %d %d %d", &kind, &n, &nsys, &largest, &parts, &pers, &w8) == 7) {
    // the ranges pw_set_pair_config accepts; other lines are skipped
    if (kind < 0 || kind > pw::kPairFP32CHO || kind == kKindLearned || parts < 1 || parts > 256) continue;
    pw::PairConfig cfg;
    cfg.parts = parts;
    cfg.persistent = pers != 0;
    if (w8 < 0 || w8 > (PW_VARIANT_WAVES8 | PW_VARIANT_SCALAR_FEED | PW_VARIANT_TWO_PER_LANE)) continue;
    const bool f32 = kind == pw::kPairFP32C || kind == pw::kPairFP32CHO;
    cfg.scalar_feed = (w8 & PW_VARIANT_SCALAR_FEED) != 0 && (f32 || kind == pw::kPairFP64);
    cfg.two_per_lane = (w8 & PW_VARIANT_TWO_PER_LANE) != 0 && cfg.scalar_feed && f32;
    cfg.waves8 = (w8 & PW_VARIANT_WAVES8) != 0 && f32 && !cfg.two_per_lane;
    tune_cache()[TuneKey(kind, n, nsys, largest)] = cfg;
  }
  std::fclose(f);
}

void save_tune_file(const TuneKey& key, const pw::PairConfig& cfg) {
  const char* path = std::getenv("PAWAN_TUNE_CACHE");
  if (!path) return;
  FILE* f = std::fopen(path, "a");
  if (!f) return;
  std::fprintf(f, "%d %lld %d %lld %d %d %d\n", std::get<0>(key), static_cast<long long>(std::get<1>(key)),
               std::get<2>(key), static_cast<long long>(std::get<3>(key)), cfg.parts,
               cfg.persistent ? 1 : 0,
               (cfg.waves8 ? PW_VARIANT_WAVES8 : 0) | (cfg.scalar_feed ? PW_VARIANT_SCALAR_FEED : 0) |
                   (cfg.two_per_lane ? PW_VARIANT_TWO_PER_LANE : 0));
  std::fclose(f);
}

int time_rhs(pw_ctx* c, int kind, double* ms, int reps = 4) {
  hipEvent_t e0, e1;
  PW_HIP(hipEventCreate(&e0));
  PW_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  for (int rep = 0; rep < reps; ++rep) {
    PW_HIP(hipEventRecord(e0, c->stream));
    const int rc = rhs(c, kind, c->y, c->k[0]);
    if (rc != PW_OK) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return rc;
    }
    PW_HIP(hipEventRecord(e1, c->stream));
    PW_HIP(hipEventSynchronize(e1));
    float t = 0.0f;
    PW_HIP(hipEventElapsedTime(&t, e0, e1));
    best = std::min(best, t);
    if (t > 50.0f) break;  // long enough that the first-launch cost is noise
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = best;
  return PW_OK;
}

}  // namespace

extern "C" {

int pw_autotune(pw_ctx* c, int precision, double* best_ms) {
  PW_CTX(c);
  int kind = 0;
  PW_TRY(kind_of(precision, &kind));
  if (kind == kKindLearned) return fail(PW_ERR_UNSUPPORTED, "launch configuration applies to the direct-sum kinds");
  PW_TRY(ensure_built(c));
  const int64_t largest = c->systems.empty() ? 0 : c->systems.front().b - c->systems.front().a;
  // keyed by the kernel family the launches run (the high-order kernels
  // have their own best shapes)
  const int family = c->smoothing == PW_KERNEL_HIGHORDER ? launch_kind_of(c, kind) : kind;
  const TuneKey key(family, c->n, c->nsys(), largest);
  auto& k = c->ks[kind];
  const bool was_profiling = c->profiling;
  c->profiling = false;
  bool found = false;
  pw::PairConfig cached;
  {
    std::lock_guard<std::mutex> lock(tune_mutex());
    load_tune_file();
    auto hit = tune_cache().find(key);
    if (hit != tune_cache().end()) {
      found = true;
      cached = hit->second;
    }
  }
  if (found) {
    k.cfg = cached;
    k.tuned = true;
    drop_graphs(c);
    // one timed right-hand side (or more, when the time is asked for), so
    // that the rates of the current state are on the device as after a
    // search
    double ms = 0.0;
    const int rc = time_rhs(c, kind, &ms, best_ms ? 4 : 1);
    c->profiling = was_profiling;
    if (rc != PW_OK) return rc;
    if (best_ms) *best_ms = ms;
    c->rates = c->k[0];
    return PW_OK;
  }
  // candidates: split the sources (fills the GPU for few target blocks and
  // cuts the tail of the last block wave), keep blocks resident, and for
  // FP32C the 8-wave register budget
  int64_t blocks = 0, max_tiles = 1;
  for (const auto& sr : c->systems) {
    const int64_t nt = (sr.b - sr.a + kTargetsPerItem - 1) / kTargetsPerItem;
    blocks += nt;
    max_tiles = std::max(max_tiles, nt);
  }
  std::vector<pw::PairConfig> cands;
  auto add = [&](int parts, bool pers, bool w8, bool sf = false, bool t2 = false) {
    if (parts > max_tiles) return;
    pw::PairConfig p;
    p.parts = parts;
    p.persistent = pers;
    p.waves8 = w8 && !t2;
    p.scalar_feed = sf;
    p.two_per_lane = t2;
    cands.push_back(p);
  };
  const bool f32 = kind == pw::kPairFP32C;
  add(1, false, false);
  add(1, true, false);
  add(2, false, false);
  add(4, true, false);
  add(8, true, false);
  if (f32) {
    add(1, false, true);
    add(4, true, true);
    add(8, true, true);
    // sources through scalar loads, one wave per 64 targets
    add(1, false, true, true);
    add(1, true, true, true);
    add(4, true, true, true);
    add(8, true, true, true);
    add(4, true, false, true);
  }
  const int64_t resident = static_cast<int64_t>(c->num_cus) * 8;
  for (int parts = 16; parts <= 64; parts *= 2)
    if (blocks * parts <= 8 * resident) add(parts, true, f32);
  // finer source ranges at any size: with the items ordered part by part,
  // a part of 1/16 or 1/32 of a large system stays in every XCD's L2 while
  // its items run
  for (int parts = 16; parts <= 32; parts *= 2) {
    if (blocks * parts > 8 * resident) add(parts, true, f32);
    if (f32) add(parts, true, true, true);
  }
  // two targets per lane (particles per thread): each scalar source record
  // feeds two pairs, at half the waves
  if (f32) {
    add(8, true, false, true, true);
    add(16, true, false, true, true);
    add(32, true, false, true, true);
  }
  // the scalar-fed FP64 kernel (Gaussian FP64 only)
  if (kind == pw::kPairFP64 && c->smoothing == PW_KERNEL_GAUSSIAN) {
    add(1, false, false, true);
    add(8, true, false, true);
    add(16, true, false, true);
    add(32, true, false, true);
  }
  pw::PairConfig best_cfg = cands.front();
  double best = 1e300;
  for (const auto& cand : cands) {
    k.cfg = cand;
    double ms = 0.0;
    const int rc = time_rhs(c, kind, &ms);
    if (rc != PW_OK) {
      c->profiling = was_profiling;
      return rc;
    }
    if (ms < best) {
      best = ms;
      best_cfg = cand;
    }
  }
  k.cfg = best_cfg;
  k.tuned = true;
  drop_graphs(c);
  {
    std::lock_guard<std::mutex> lock(tune_mutex());
    tune_cache()[key] = best_cfg;
    save_tune_file(key, best_cfg);
  }
  // the rates of the chosen configuration (the candidates' reductions of
  // source-split partial sums differ in rounding)
  {
    const int rc = rhs(c, kind, c->y, c->k[0]);
    c->profiling = was_profiling;
    if (rc != PW_OK) return rc;
  }
  c->rates = c->k[0];
  if (best_ms) *best_ms = best;
  return PW_OK;
}

}  // extern "C"
