// axb_capi.hip -- the extern "C" boundary (include/arborx_mi355x.h).
//
// Orchestrates the kernels of axb_build.hip / axb_sort.hip / axb_query.hip /
// axb_energy.hip on the caller's stream, mirroring the reference's host-side
// control flow (spatial/ArborX_LinearBVH.hpp:167-313,
// detail/ArborX_CrsGraphWrapperImpl.hpp:145-466).  Host synchronisation
// happens where the reference has one: after the build (bounds(),
// TreeConstruction.hpp:108-113) and after the counting pass of a spatial
// query (lastElement, CrsGraphWrapperImpl.hpp:248).
#include "arborx_mi355x.h"
#include "axb_internal.hpp"

#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <array>
#include <map>
#include <chrono>
#include <atomic>
#include <vector>
#include <cstring>
#include <mutex>
#include <string>

namespace axb
{
namespace
{
thread_local std::string g_error;
std::mutex g_alloc_mutex;
axb_alloc_fn g_alloc = nullptr;
axb_free_fn g_free = nullptr;
void *g_alloc_ctx = nullptr;
} // namespace

void set_error(const std::string &msg) { g_error = msg; }

// ------------------------------------------------------------ profiling
namespace
{
struct ProfRec
{
  std::string name;
  hipEvent_t a, b;
  double host_ms; // >= 0: a host-side interval (no events)
};
std::atomic<bool> g_prof_timing{false};
std::atomic<bool> g_prof_visits{false};
std::mutex g_prof_mutex;
std::vector<ProfRec> g_prof_recs;
unsigned long long *g_visits = nullptr; // device, 4 counters
} // namespace

ProfScope::ProfScope(const char *name, hipStream_t s) : name_(name), s_(s), on_(g_prof_timing.load())
{
  if (on_)
  {
    on_ = hipEventCreate(&a_) == hipSuccess && hipEventCreate(&b_) == hipSuccess &&
          hipEventRecord(a_, s_) == hipSuccess;
  }
}
ProfScope::~ProfScope()
{
  if (on_ && hipEventRecord(b_, s_) == hipSuccess)
  {
    std::lock_guard<std::mutex> lock(g_prof_mutex);
    g_prof_recs.push_back({name_, a_, b_, -1.0});
  }
}

Region::Region(const char *label) { roctxRangePushA(label); }
Region::~Region() { roctxRangePop(); }
void Region::next(const char *label)
{
  roctxRangePop();
  roctxRangePushA(label);
}

HostScope::HostScope(const char *name) : name_(name), on_(g_prof_timing.load())
{
  if (on_)
    t0_ = std::chrono::steady_clock::now();
}
HostScope::~HostScope()
{
  if (on_)
  {
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
    std::lock_guard<std::mutex> lock(g_prof_mutex);
    g_prof_recs.push_back({name_, nullptr, nullptr, ms});
  }
}

hipError_t sync_stream(hipStream_t s)
{
  HostScope h("host:stream_sync");
  return hipStreamSynchronize(s);
}
unsigned long long *visit_counters() { return g_prof_visits.load() ? g_visits : nullptr; }

// Temporaries come from the device's stream-ordered pool.  By default the
// pool returns freed memory to the driver at every synchronisation, which
// costs milliseconds per call at these sizes; the library keeps it cached
// instead (once per device, on first use).
static void keep_pool_memory()
{
  static std::mutex m;
  static std::vector<bool> done;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
  {
    (void)hipGetLastError();
    return;
  }
  std::lock_guard<std::mutex> lock(m);
  if ((int)done.size() <= dev)
    done.resize(dev + 1, false);
  if (done[dev])
    return;
  done[dev] = true;
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess)
  {
    uint64_t keep = UINT64_MAX;
    if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess)
      (void)hipGetLastError();
  }
  else
    (void)hipGetLastError();
}

void *dalloc(size_t bytes, hipStream_t stream)
{
  HostScope h("host:dalloc");
  keep_pool_memory();
  void *p = nullptr;
  if (bytes == 0)
    bytes = 16;
  if (hipMallocAsync(&p, bytes, stream) != hipSuccess)
  {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}
void dfree(void *p, hipStream_t stream)
{
  HostScope h("host:dfree");
  if (p)
    (void)hipFreeAsync(p, stream);
}

// Grow-only workspaces per (device, stream, slot) for the large per-call
// scratch buffers (scratch rows of the radius CRS, kNN candidate rows):
// freeing buffers of gigabytes costs milliseconds per call, so they are
// kept for the next call on the same stream (stream order keeps the uses
// apart).  axb_trim_memory releases them.
namespace
{
struct WsKey
{
  int dev;
  hipStream_t s;
  bool operator<(const WsKey &o) const { return dev < o.dev || (dev == o.dev && s < o.s); }
};
struct WsBlock
{
  void *p = nullptr;
  size_t bytes = 0;
  bool busy = false; // leased to a call
};
std::mutex g_ws_mutex;
std::map<WsKey, std::array<WsBlock, WS_NUM>> g_ws;

std::atomic<int64_t> g_lim_knn_batch{0};
std::atomic<int64_t> g_lim_rows_bytes{0};
std::atomic<int32_t> g_lim_knn_rounds{0};
} // namespace

Limits debug_limits()
{
  return Limits{g_lim_knn_batch.load(), g_lim_rows_bytes.load(), g_lim_knn_rounds.load()};
}

void WsLease::acquire(hipStream_t s_, int slot_, size_t bytes)
{
  HostScope h("host:ws_get");
  release();
  s = s_;
  slot = slot_;
  if (hipGetDevice(&dev) != hipSuccess)
  {
    (void)hipGetLastError();
    return;
  }
  {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    WsBlock &b = g_ws[WsKey{dev, s}][slot];
    if (!b.busy)
    {
      // the block's earlier users ran on this same stream, so a free
      // enqueued here follows their kernels
      if (!(b.p && b.bytes >= bytes))
      {
        if (b.p)
          dfree(b.p, s);
        b.p = dalloc(bytes, s);
        b.bytes = b.p ? bytes : 0;
      }
      if (b.p)
      {
        b.busy = true;
        cached = true;
        p = b.p;
      }
      return;
    }
  }
  p = dalloc(bytes, s);
  cached = false;
}

void WsLease::release()
{
  if (!p)
    return;
  if (cached)
  {
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    auto it = g_ws.find(WsKey{dev, s});
    if (it != g_ws.end())
      it->second[slot].busy = false;
  }
  else
    dfree(p, s);
  p = nullptr;
  cached = false;
}

// Releases every cached workspace not leased to a call (after the device
// is idle).
static void ws_release_all()
{
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  for (auto &kv : g_ws)
    for (WsBlock &b : kv.second)
      if (b.p && !b.busy)
      {
        (void)hipSetDevice(kv.first.dev);
        (void)hipFree(b.p);
        b = WsBlock{};
      }
}

// Result arrays the caller owns.
static void *out_alloc(size_t bytes, hipStream_t stream)
{
  HostScope h("host:out_alloc");
  std::lock_guard<std::mutex> lock(g_alloc_mutex);
  if (g_alloc)
    return g_alloc(bytes == 0 ? 16 : bytes, g_alloc_ctx);
  return dalloc(bytes, stream);
}
static void out_free(void *p, hipStream_t stream)
{
  if (!p)
    return;
  std::lock_guard<std::mutex> lock(g_alloc_mutex);
  if (g_free)
    g_free(p, g_alloc_ctx);
  else if (!g_alloc)
    dfree(p, stream);
}

} // namespace axb

using namespace axb;

struct axb_bvh
{
  int64_t n = 0;
  int geometry = 0;
  int node_format = 0;
  bool keep_topology = false;
  BuildOut b{};
  float bounds[6];
  hipStream_t stream = nullptr;
  uint32_t *d_err = nullptr;
};

#define AXB_FAIL(code, msg)                                                                       \
  do                                                                                              \
  {                                                                                               \
    ::axb::set_error(msg);                                                                        \
    return code;                                                                                  \
  } while (0)

static axb_status hip_fail(hipError_t e, const char *where)
{
  set_error(std::string(where) + ": " + hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? AXB_ERR_OUT_OF_MEMORY : AXB_ERR_HIP;
}

// Reads and clears a kernel error word (the stream must be synchronised).
// `tolerated` (optional): a traversal-stack overflow alone is reported there
// instead of as an error, for callers that redo the work another way.
// kernel_error_word: the same for an error word already copied to the host.
static axb_status kernel_error_word(uint32_t h, uint32_t *d_err, hipStream_t s,
                                    const char *where, uint32_t *tolerated = nullptr)
{
  if (h && tolerated && h == KERR_STACK)
  {
    (void)fill_u32(d_err, 0u, 1, s);
    *tolerated = h;
    return AXB_OK;
  }
  if (h)
  {
    (void)fill_u32(d_err, 0u, 1, s);
    std::string m = std::string(where) + ": kernel reported";
    if (h & KERR_SPIN)
      m += " bounded-wait timeout";
    if (h & KERR_STACK)
      m += " traversal stack overflow";
    if (h & KERR_FRAME)
      m += " Q16 box outside its frame";
    if (h & KERR_LAYOUT)
      m += " inconsistent node layout";
    set_error(m);
    return AXB_ERR_INTERNAL;
  }
  return AXB_OK;
}

static axb_status check_kernel_error(uint32_t *d_err, hipStream_t s, const char *where,
                                     uint32_t *tolerated = nullptr)
{
  uint32_t h = 0;
  hipError_t e = hipMemcpyAsync(&h, d_err, sizeof(h), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = sync_stream(s);
  if (e != hipSuccess)
    return hip_fail(e, where);
  return kernel_error_word(h, d_err, s, where, tolerated);
}

// The kernel error word of one call: calls running at the same time on
// other streams never see (or clear) each other's flags.
struct CallErr
{
  uint32_t *p = nullptr;
  hipStream_t s;
  explicit CallErr(hipStream_t s_) : s(s_)
  {
    p = (uint32_t *)dalloc(sizeof(uint32_t), s);
    if (p && fill_u32(p, 0u, 1, s) != hipSuccess)
    {
      dfree(p, s);
      p = nullptr;
    }
  }
  CallErr(const CallErr &) = delete;
  CallErr &operator=(const CallErr &) = delete;
  ~CallErr() { dfree(p, s); }
};

// Device view of a tree for one call's kernels (err: that call's word).
static TreeView view_of(const axb_bvh *t, uint32_t *err)
{
  TreeView tv;
  tv.slots = t->b.slots;
  tv.total = (uint32_t)t->b.total_slots;
  tv.n = (int32_t)t->n;
  tv.C = t->b.C;
  tv.err = err;
  tv.visits = visit_counters();
  for (int d = 0; d < 3; ++d)
  {
    tv.scene.lo[d] = t->bounds[d];
    tv.scene.hi[d] = t->bounds[3 + d];
    tv.ss[d] = q16_scale(t->bounds[d], t->bounds[3 + d]);
  }
  return tv;
}

extern "C" {

const char *axb_last_error(void) { return g_error.c_str(); }

void axb_q16_encode_host(const float frame[6], const float box[6], uint16_t q[6])
{
  for (int d = 0; d < 3; ++d)
  {
    uint32_t lo = 0, hi = 0;
    q16_encode_axis(frame[d], frame[3 + d], box[d], box[3 + d], lo, hi);
    q[d] = (uint16_t)lo;
    q[3 + d] = (uint16_t)hi;
  }
}

axb_status axb_trim_memory(void)
{
  int dev = 0;
  AXB_HIP_TRY(hipGetDevice(&dev));
  AXB_HIP_TRY(hipDeviceSynchronize());
  ws_release_all();
  AXB_HIP_TRY(hipSetDevice(dev));
  hipMemPool_t pool;
  AXB_HIP_TRY(hipDeviceGetDefaultMemPool(&pool, dev));
  AXB_HIP_TRY(hipMemPoolTrimTo(pool, 0));
  return AXB_OK;
}
const char *axb_version(void) { return "arborx-mi355x 0.1 (gfx950)"; }

axb_status axb_set_output_allocator(axb_alloc_fn alloc, axb_free_fn free_fn, void *ctx)
{
  if ((alloc == nullptr) != (free_fn == nullptr))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_set_output_allocator: give both functions or none");
  std::lock_guard<std::mutex> lock(g_alloc_mutex);
  g_alloc = alloc;
  g_free = free_fn;
  g_alloc_ctx = ctx;
  return AXB_OK;
}

axb_status axb_free(void *device_ptr, axb_stream_t stream)
{
  if (device_ptr)
  {
    hipError_t e = hipFreeAsync(device_ptr, (hipStream_t)stream);
    if (e != hipSuccess)
      return hip_fail(e, "axb_free");
  }
  return AXB_OK;
}

float axb_sqrt_threshold(float r) { return sqrt_threshold(r); }

axb_status axb_debug_set_limits(int64_t knn_batch_max, int64_t scratch_rows_bytes,
                                int32_t knn_rounds)
{
  if (knn_batch_max < 0 || scratch_rows_bytes < 0 || knn_rounds < -1 || knn_rounds > 64)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_debug_set_limits: bad limit");
  g_lim_knn_batch.store(knn_batch_max);
  g_lim_rows_bytes.store(scratch_rows_bytes);
  g_lim_knn_rounds.store(knn_rounds);
  return AXB_OK;
}

uint64_t axb_morton64_host(const float p[3], const float scene[6])
{
  return morton64(translate_and_scale(p[0], scene[0], scene[3]),
                  translate_and_scale(p[1], scene[1], scene[4]),
                  translate_and_scale(p[2], scene[2], scene[5]));
}

axb_status axb_bvh_create(const float *values, int64_t n, int32_t geometry, int32_t node_format,
                          int32_t flags, axb_stream_t stream_, axb_bvh **out)
{
  Region region("ArborX::BVH::BVH");
  if (!out)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_create: out is NULL");
  *out = nullptr;
  if (n < 0 || n >= (int64_t(1) << 30))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_create: n must be in [0, 2^30)");
  if (geometry != AXB_GEOMETRY_POINTS && geometry != AXB_GEOMETRY_BOXES)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_create: unknown geometry");
  if (node_format != AXB_NODES_FP32 && node_format != AXB_NODES_Q16)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_create: unknown node format");
  if (n > 0 && !values)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_create: values is NULL");
  hipStream_t s = (hipStream_t)stream_;
  axb_bvh *t = new axb_bvh;
  t->n = n;
  t->geometry = geometry;
  t->node_format = node_format;
  t->keep_topology = (flags & AXB_BUILD_KEEP_TOPOLOGY) != 0;
  t->stream = s;
  // error word, then the build's summary (stored internal nodes, scene box)
  t->d_err = (uint32_t *)dalloc(BUILD_STATUS_WORDS * sizeof(uint32_t), s);
  if (!t->d_err)
  {
    delete t;
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "axb_bvh_create: out of device memory");
  }
  if (fill_u32(t->d_err, 0u, 1, s) != hipSuccess)
  {
    dfree(t->d_err, s);
    delete t;
    AXB_FAIL(AXB_ERR_HIP, "axb_bvh_create: kernel launch failed");
  }
  hipError_t e = build_tree(values, n, geometry, node_format, t->keep_topology, t->d_err, s, t->b);
  // one copy to the host: error word, stored internal nodes (C[n - 1]), scene
  uint32_t status[BUILD_STATUS_WORDS] = {};
  if (e == hipSuccess)
    e = build_status(t->b, n, t->d_err, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(status, t->d_err, sizeof(status), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = sync_stream(s);
  const int32_t c_last = (int32_t)status[1];
  std::memcpy(t->bounds, status + 2, sizeof(t->bounds));
  if (e == hipSuccess && n > 0)
  {
    const int64_t sl = geometry == AXB_GEOMETRY_POINTS ? 1 : 2;
    const int64_t si = node_format == AXB_NODES_FP32 ? 2 : 1;
    const int64_t total = sl * n + si * (int64_t)c_last;
    if (c_last < 0 || total > t->b.max_slots)
      e = hipErrorUnknown;
    else
      t->b.total_slots = total;
  }
  if (e != hipSuccess)
  {
    axb_status st = hip_fail(e, "axb_bvh_create");
    axb_bvh_destroy(t);
    return st;
  }
  axb_status st = kernel_error_word(status[0], t->d_err, s, "axb_bvh_create");
  if (st != AXB_OK)
  {
    axb_bvh_destroy(t);
    return st;
  }
  if (node_format == AXB_NODES_Q16 && n > 1)
  {
    for (int d = 0; d < 3; ++d)
      if (!std::isfinite(t->bounds[3 + d] - t->bounds[d]))
      {
        axb_bvh_destroy(t);
        AXB_FAIL(AXB_ERR_INVALID_ARGUMENT,
                 "axb_bvh_create: Q16 nodes need a scene box of finite extent");
      }
  }
  *out = t;
  return AXB_OK;
}

void axb_bvh_destroy(axb_bvh *t)
{
  if (!t)
    return;
  hipStream_t s = t->stream;
  dfree(t->b.slots, s);
  dfree(t->b.scene, s);
  dfree(t->b.C, s);
  dfree(t->b.keys, s);
  dfree(t->b.perm, s);
  dfree(t->b.leaf_rope, s);
  dfree(t->b.inodes, s);
  dfree(t->d_err, s);
  delete t;
}

int64_t axb_bvh_size(const axb_bvh *t) { return t ? t->n : 0; }

axb_status axb_bvh_bounds(const axb_bvh *t, float out_host[6])
{
  if (!t || !out_host)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_bounds: NULL argument");
  std::memcpy(out_host, t->bounds, sizeof(t->bounds));
  return AXB_OK;
}

int64_t axb_bvh_num_slots(const axb_bvh *t) { return t ? t->b.total_slots : 0; }
const void *axb_bvh_slots(const axb_bvh *t) { return t ? t->b.slots : nullptr; }

axb_status axb_bvh_export_topology(const axb_bvh *t, uint64_t *keys, uint32_t *perm,
                                   int32_t *leaf_rope, int32_t *left_child, int32_t *rope,
                                   float *boxes, axb_stream_t stream_)
{
  if (!t)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_export_topology: NULL tree");
  if (t->n < 2)
    return AXB_OK;
  if (!t->keep_topology)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT,
             "axb_bvh_export_topology: tree was built without AXB_BUILD_KEEP_TOPOLOGY");
  hipStream_t s = (hipStream_t)stream_;
  const size_t n = (size_t)t->n;
  if (keys)
    AXB_HIP_TRY(hipMemcpyAsync(keys, t->b.keys, n * 8, hipMemcpyDeviceToDevice, s));
  if (perm)
    AXB_HIP_TRY(hipMemcpyAsync(perm, t->b.perm, n * 4, hipMemcpyDeviceToDevice, s));
  if (leaf_rope)
    AXB_HIP_TRY(hipMemcpyAsync(leaf_rope, t->b.leaf_rope, n * 4, hipMemcpyDeviceToDevice, s));
  AXB_HIP_TRY(export_topology(t->b.inodes, t->n - 1, left_child, rope, boxes, s));
  return AXB_OK;
}

// Morton-32 processing order of predicates (CrsGraphWrapperImpl.hpp:407-419).
// exact: the permutation that sorts the centroids by their 30-bit Morton
// codes (ties by position); otherwise by the top 24 bits of the codes (8 per
// axis, cells of 1/256 of the scene per axis), one radix pass fewer -- the
// processing order of a query batch only decides which predicates share a
// wavefront, never a result.
static axb_status morton_order(const float *scene, uint32_t *d_err, int pred_kind,
                               const float *preds, int64_t nq, hipStream_t s,
                               uint32_t **order_out, void **to_free, bool exact = false)
{
  *order_out = nullptr;
  if (nq < 2)
    return AXB_OK;
  // the sort's look-back words carry 30-bit counts
  if (nq >= MAX_SORTED_PREDICATES)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "query: sort_predicates needs fewer than 2^30 predicates");
  const int passes = exact ? 4 : 3;
  const size_t ws_bytes = sort_workspace_bytes(nq, 4, 4);
  // one block: keys0 keys1 vals0 vals1 hist workspace
  const size_t a = (size_t)nq * 4;
  const size_t off_k1 = a, off_v0 = 2 * a, off_v1 = 3 * a, off_h = 4 * a,
               off_ws = 4 * a + 4 * 256 * 4;
  char *blk = (char *)dalloc(off_ws + ws_bytes, s);
  if (!blk)
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "query: out of device memory (predicate sort)");
  uint32_t *k0 = (uint32_t *)blk, *k1 = (uint32_t *)(blk + off_k1);
  uint32_t *v0 = (uint32_t *)(blk + off_v0), *v1 = (uint32_t *)(blk + off_v1);
  uint32_t *hist = (uint32_t *)(blk + off_h);
  hipError_t e = fill_u32(hist, 0u, 4 * 256, s);
  if (e == hipSuccess)
    e = launch_pred_morton32(pred_kind, preds, nq, scene, exact ? 0 : 6, k0, hist, s);
  if (e == hipSuccess)
    e = radix_sort_pairs_u32(k0, k1, v0, v1, nq, passes, hist, blk + off_ws, d_err, s);
  if (e != hipSuccess)
  {
    dfree(blk, s);
    return hip_fail(e, "query: predicate sort");
  }
  // each pass moves the pairs to the other buffer
  *order_out = passes % 2 == 0 ? v0 : v1;
  *to_free = blk;
  return AXB_OK;
}

static axb_status predicate_order(const axb_bvh *t, uint32_t *err, int pred_kind,
                                  const float *preds, int64_t nq, hipStream_t s,
                                  uint32_t **order_out, void **to_free, bool exact = false)
{
  return morton_order(t->b.scene, err, pred_kind, preds, nq, s, order_out, to_free, exact);
}

} // extern "C"

// Bytes the scratch rows of a buffered CRS query may take: an eighth of the
// current device's memory (36 GB on MI355X: 96 results per row for up to
// 93M queries, shorter rows beyond), at least 4 GiB.
static int64_t scratch_rows_budget()
{
  static std::atomic<int64_t> budget[64] = {};
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::atomic<int64_t> &b = budget[dev & 63];
  const int64_t lim = debug_limits().scratch_rows_bytes;
  if (lim > 0)
    return lim;
  int64_t v = b.load(std::memory_order_relaxed);
  if (v == 0)
  {
    size_t total = 0;
    v = hipDeviceTotalMem(&total, dev) == hipSuccess ? (int64_t)(total / 8) : 0;
    v = std::max<int64_t>(v, int64_t(4) << 30);
    b.store(v, std::memory_order_relaxed);
  }
  return v;
}

// One CRS query of nq rows (CrsGraphWrapperImpl.hpp:145-374 semantics).
// first(mode, out, B) traverses every row once: MODE_COUNT, or MODE_BUFFER
// keeping up to B results per row in a scratch row.  Counts become offsets
// by an exclusive scan; the rows that fit are compacted, and only the rows
// that overflowed are traversed again by fill(list, n_list, out) with
// MODE_FILL (list == nullptr: every row).  The overflow list walks the
// processing order `ovf_order` (processing index -> row, nullptr: identity)
// and holds rows, or processing indices when `ovf_emit_index`.
// TraversalPolicy buffer_size: 0 -> a scratch row of up to `auto_rows`
// results (96 for the tree's queries and neighbour lists, 64 for brute
// force) when the scratch fits scratch_rows_budget(),
// else plain count + fill; > 0 -> that many;
// < 0 -> that many, and a row that overflows is an error
// (AXB_ERR_BUFFER_OVERFLOW, the reference's SearchException).
template <class First, class Fill>
static axb_status crs_query(uint32_t *d_err, int64_t nq, axb_traversal_policy policy,
                            const uint32_t *ovf_order, bool ovf_emit_index, First first,
                            Fill fill, int64_t *offsets, int32_t **indices, int64_t *nnz_host,
                            hipStream_t s, const char *what, uint32_t *tolerated = nullptr,
                            int64_t auto_rows = 64)
{
  Region region("ArborX::CrsGraphWrapper::two_pass");
  const bool hard = policy.buffer_size < 0;
  int64_t B = policy.buffer_size < 0 ? -(int64_t)policy.buffer_size : policy.buffer_size;
  if (B == 0)
  {
    const int64_t fit = (scratch_rows_budget() / (4 * nq)) & ~int64_t(3);
    B = fit >= auto_rows ? auto_rows : (fit >= 16 ? fit : 0);
  }
  void *scan_ws = dalloc(scan_workspace_bytes(nq + 1), s);
  WsLease rows;
  int32_t *buf = nullptr;
  uint32_t *ovf = nullptr, *ovf_n = nullptr;
  uint8_t *over_flags = nullptr; // a bit per row: its count exceeds B
  axb_status st = AXB_OK;
  hipError_t e = scan_ws ? hipSuccess : hipErrorOutOfMemory;
  int64_t nnz = 0;
  uint32_t n_ovf = 0;
  if (e == hipSuccess && B > 0)
  {
    rows.acquire(s, WS_ROWS, (size_t)nq * B * sizeof(int32_t));
    buf = (int32_t *)rows.p;
    ovf = (uint32_t *)dalloc((size_t)nq * sizeof(uint32_t), s);
    ovf_n = (uint32_t *)dalloc(sizeof(uint32_t), s);
    over_flags = (uint8_t *)dalloc((size_t)div_up(nq + 1, 8), s);
    e = (buf && ovf && ovf_n && over_flags) ? hipSuccess : hipErrorOutOfMemory;
    if (e == hipSuccess)
      e = first(MODE_BUFFER, buf, (int32_t)B);
  }
  else if (e == hipSuccess)
    e = first(MODE_COUNT, nullptr, 0);
  // counts -> offsets (exclusive scan over nq + 1 entries, last count 0)
  if (e == hipSuccess)
    e = fill_u64((uint64_t *)(offsets + nq), 0ull, 1, s);
  if (e == hipSuccess)
    e = exclusive_scan_i64(offsets, offsets, nq + 1, scan_ws, d_err, s, over_flags, B);
  if (e == hipSuccess && B > 0)
  {
    e = fill_u32(ovf_n, 0u, 1, s);
    if (e == hipSuccess)
      e = launch_overflow_list(over_flags, ovf_order, nq, ovf_emit_index, ovf, ovf_n, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(&n_ovf, ovf_n, sizeof(n_ovf), hipMemcpyDeviceToHost, s);
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(&nnz, offsets + nq, sizeof(nnz), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = sync_stream(s);
  if (e == hipSuccess && hard && n_ovf > 0)
  {
    set_error(std::string(what) + ": a row has more results than the hard buffer size "
                                  "(TraversalPolicy buffer_size < 0)");
    st = AXB_ERR_BUFFER_OVERFLOW;
  }
  if (e == hipSuccess && st == AXB_OK)
  {
    *indices = (int32_t *)out_alloc((size_t)nnz * sizeof(int32_t), s);
    if (!*indices)
      e = hipErrorOutOfMemory;
  }
  if (e == hipSuccess && st == AXB_OK && nnz > 0)
  {
    if (B == 0)
      e = fill(nullptr, nq, *indices);
    else
    {
      e = launch_compact(buf, nq, (int32_t)B, offsets, *indices, s);
      // rows that did not fit: traverse those again, writing in place
      if (e == hipSuccess && n_ovf > 0)
        e = fill(ovf, n_ovf, *indices);
    }
  }
  if (e == hipSuccess)
    e = sync_stream(s);
  dfree(scan_ws, s);
  dfree(ovf, s);
  dfree(ovf_n, s);
  dfree(over_flags, s);
  if (e != hipSuccess)
  {
    out_free(*indices, s);
    *indices = nullptr;
    return hip_fail(e, what);
  }
  if (st != AXB_OK)
    return st;
  st = check_kernel_error(d_err, s, what, tolerated);
  if (st != AXB_OK)
  {
    out_free(*indices, s);
    *indices = nullptr;
    return st;
  }
  *nnz_host = nnz;
  return AXB_OK;
}

// The geometry kind of the device's ordered walk for an ordered predicate
// kind of the C ABI (-1: not one of them).
static int ordered_geom_kind(int pred_kind)
{
  switch (pred_kind)
  {
  case AXB_PREDICATE_ORDERED_SPHERES:
    return PRED_SPHERES;
  case AXB_PREDICATE_ORDERED_BOXES:
    return PRED_BOXES;
  case AXB_PREDICATE_ORDERED_POINTS:
    return PRED_POINTS;
  default:
    return -1;
  }
}

// CRS of OrderedSpatial<Sphere | Box | Point>: every row lists all n values
// by (distance, index), offsets[q] = q n.  The traversal policy's buffer
// size only matters when negative (an error if n exceeds it, as the
// reference's overflow check); sorting predicates changes nothing.
static axb_status query_ordered_rows(const axb_bvh *t, int geom_kind, const float *preds,
                                     int64_t nq, axb_traversal_policy policy, int64_t *offsets,
                                     int32_t **indices, int64_t *nnz_host, hipStream_t s)
{
  const int64_t n = t->n;
  if (policy.buffer_size < 0 && n > -(int64_t)policy.buffer_size)
    AXB_FAIL(AXB_ERR_BUFFER_OVERFLOW, "query: a row has more results than the hard buffer "
                                      "size (TraversalPolicy buffer_size < 0)");
  CallErr err(s);
  if (!err.p)
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "query: out of device memory");
  const TreeView tv = view_of(t, err.p);
  const int64_t nnz = nq * n;
  hipError_t e = launch_knn_offsets(offsets, nq, n, s);
  if (e == hipSuccess)
  {
    *indices = (int32_t *)out_alloc((size_t)nnz * sizeof(int32_t), s);
    if (!*indices)
      e = hipErrorOutOfMemory;
  }
  if (e == hipSuccess)
    e = launch_ordered_rows(tv, t->node_format, t->geometry, geom_kind, preds, nq, *indices, s);
  if (e == hipSuccess)
    e = sync_stream(s);
  if (e != hipSuccess)
  {
    out_free(*indices, s);
    *indices = nullptr;
    return hip_fail(e, "query: ordered rows");
  }
  const axb_status st = check_kernel_error(err.p, s, "query: ordered rows");
  if (st != AXB_OK)
  {
    out_free(*indices, s);
    *indices = nullptr;
    return st;
  }
  *nnz_host = nnz;
  return AXB_OK;
}

extern "C" {

// Spatial CRS queries; pred_kind AXB_PREDICATE_ORDERED_RAYS takes the
// best-first walk, rows cut at `limit` results (Nearest<Ray>: k).
static axb_status query_spatial(const axb_bvh *t, int pred_kind, const float *preds, int64_t nq,
                                float radius, axb_traversal_policy policy, int64_t *offsets,
                                int32_t **indices, int64_t *nnz_host, axb_stream_t stream_,
                                int32_t limit = INT32_MAX, uint32_t *tolerated = nullptr)
{
  Region region("ArborX::BVH::query_crs");
  Region tag_region(pred_kind == AXB_PREDICATE_ORDERED_RAYS || ordered_geom_kind(pred_kind) >= 0
                        ? "ArborX::BVH::query::ordered_spatial"
                        : "ArborX::BVH::query::spatial");
  if (!t || !offsets || !indices || !nnz_host)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "query: NULL argument");
  const bool ordered = pred_kind == AXB_PREDICATE_ORDERED_RAYS;
  if ((ordered || pred_kind == PRED_RAYS) && t->geometry != GEOM_BOXES && t->n > 0)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "query: ray predicates need a tree of boxes");
  if (nq < 0 || nq >= (int64_t(1) << 31))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "query: nq must be in [0, 2^31)");
  if (nq > 0 && !preds)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "query: predicates is NULL");
  if (policy.sort_predicates && nq >= MAX_SORTED_PREDICATES)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "query: sort_predicates needs fewer than 2^30 predicates");
  hipStream_t s = (hipStream_t)stream_;
  *indices = nullptr;
  *nnz_host = 0;
  if (nq == 0 || t->n == 0)
  {
    AXB_HIP_TRY(fill_u64((uint64_t *)offsets, 0ull, nq + 1, s));
    *indices = (int32_t *)out_alloc(0, s);
    AXB_HIP_TRY(sync_stream(s));
    return AXB_OK;
  }
  const int geom_kind = ordered_geom_kind(pred_kind);
  if (geom_kind >= 0)
    return query_ordered_rows(t, geom_kind, preds, nq, policy, offsets, indices, nnz_host, s);
  CallErr err(s);
  if (!err.p)
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "query: out of device memory");
  const TreeView tv = view_of(t, err.p);
  uint32_t *order = nullptr;
  void *order_blk = nullptr;
  if (policy.sort_predicates)
  {
    Region perm_region(ordered ? "ArborX::BVH::query::ordered_spatial::compute_permutation"
                               : "ArborX::BVH::query::spatial::compute_permutation");
    axb_status st = predicate_order(t, err.p, ordered ? PRED_RAYS : pred_kind, preds, nq, s,
                                    &order, &order_blk);
    if (st != AXB_OK)
      return st;
  }
  // whole ordered rows: the rope walk collects the slots of the hit leaves
  // and each row is then sorted by (distance, index); the best-first walk
  // is kept for rows cut at `limit` (it stops early)
  const bool sorted_rows = ordered && limit == INT32_MAX;
  auto first = [&](int mode, int32_t *out, int32_t B) {
    if (sorted_rows)
      return launch_ray_slots(tv, t->node_format, preds, nq, order, offsets, out, B, mode, s);
    if (ordered)
      return launch_ordered(tv, t->node_format, preds, nq, limit, order, offsets, out, B, mode,
                            s);
    return launch_spatial(tv, t->node_format, t->geometry, pred_kind, preds, nq, radius, order,
                          offsets, out, B, mode, s);
  };
  // the overflow list holds rows; the fill pass takes them as its order
  auto fill = [&](const uint32_t *list, int64_t n_list, int32_t *out) {
    if (sorted_rows)
      return launch_ray_slots(tv, t->node_format, preds, n_list, list ? list : order, offsets,
                              out, 0, MODE_FILL, s);
    if (ordered)
      return launch_ordered(tv, t->node_format, preds, n_list, limit, list ? list : order,
                            offsets, out, 0, MODE_FILL, s);
    return launch_spatial(tv, t->node_format, t->geometry, pred_kind, preds, n_list, radius,
                          list ? list : order, offsets, out, 0, MODE_FILL, s);
  };
  // Rows of 96: rows that overflow their scratch row are walked again, and
  // row counts have tails -- rays (1M rays over 5M boxes: mean 10, 0.1%
  // above 64; intersects 21.2 -> 16.1 ms against rows of 64), radius rows
  // of clustered points (10M-point Gaussian mixture, ~32 neighbours: 8.6 ->
  // 7.1 ms) -- while rows of ~32 uniform neighbours take the same time with
  // 64 or 96 (6.9 ms at 10M points).
  axb_status st = crs_query(err.p, nq, policy, order, false, first, fill, offsets, indices,
                            nnz_host, s, "query", tolerated, 96);
  dfree(order_blk, s);
  if (st == AXB_OK && sorted_rows && *nnz_host > 0)
  {
    hipError_t e = launch_ray_sort_rows(tv, preds, nq, offsets, *indices, s);
    if (e == hipSuccess)
      e = sync_stream(s);
    if (e != hipSuccess)
    {
      out_free(*indices, s);
      *indices = nullptr;
      return hip_fail(e, "query: ordered rows");
    }
  }
  return st;
}

axb_status axb_bvh_query_intersects(const axb_bvh *t, int32_t predicate_kind,
                                    const float *predicates, int64_t nq,
                                    axb_traversal_policy policy, int64_t *offsets,
                                    int32_t **indices, int64_t *nnz_host, axb_stream_t stream)
{
  if (predicate_kind != AXB_PREDICATE_SPHERES && predicate_kind != AXB_PREDICATE_BOXES &&
      predicate_kind != AXB_PREDICATE_RAYS && predicate_kind != AXB_PREDICATE_ORDERED_RAYS &&
      ordered_geom_kind(predicate_kind) < 0)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_query_intersects: unknown predicate kind");
  return query_spatial(t, predicate_kind, predicates, nq, 0.f, policy, offsets, indices,
                       nnz_host, stream);
}

axb_status axb_bvh_query_nearest_rays(const axb_bvh *t, const float *rays, int64_t nq,
                                      int32_t k, axb_traversal_policy policy,
                                      int64_t *offsets, int32_t **indices, int64_t *nnz_host,
                                      axb_stream_t stream)
{
  // TreeTraversal.hpp:143-178: k < 1 returns nothing
  uint32_t tol = 0;
  axb_status st = query_spatial(t, AXB_PREDICATE_ORDERED_RAYS, rays, nq, 0.f, policy, offsets,
                                indices, nnz_host, stream, k > 0 ? k : 0, &tol);
  if (st != AXB_OK || !(tol & KERR_STACK))
    return st;
  // A ray's best-first walk outgrew its heap: the rows are then taken as
  // the first k of the whole ordered rows (rope walk + per-row sort).
  hipStream_t s = (hipStream_t)stream;
  out_free(*indices, s);
  *indices = nullptr;
  *nnz_host = 0;
  int64_t *full_off = (int64_t *)dalloc((size_t)(nq + 1) * sizeof(int64_t), s);
  void *scan_ws = dalloc(scan_workspace_bytes(nq + 1), s);
  if (!full_off || !scan_ws)
  {
    dfree(full_off, s);
    dfree(scan_ws, s);
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "axb_bvh_query_nearest_rays: out of device memory");
  }
  CallErr err(s);
  if (!err.p)
  {
    dfree(full_off, s);
    dfree(scan_ws, s);
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "axb_bvh_query_nearest_rays: out of device memory");
  }
  int32_t *full_ind = nullptr;
  int64_t full_nnz = 0;
  axb_traversal_policy whole = policy;
  whole.buffer_size = 0; // whole rows may be longer than k
  st = query_spatial(t, AXB_PREDICATE_ORDERED_RAYS, rays, nq, 0.f, whole, full_off, &full_ind,
                     &full_nnz, stream);
  hipError_t e = hipSuccess;
  int64_t nnz = 0;
  if (st == AXB_OK)
  {
    e = launch_row_prefix_counts(full_off, nq, k, offsets, s);
    if (e == hipSuccess)
      e = exclusive_scan_i64(offsets, offsets, nq + 1, scan_ws, err.p, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(&nnz, offsets + nq, sizeof(nnz), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
      e = sync_stream(s);
    if (e == hipSuccess)
    {
      *indices = (int32_t *)out_alloc((size_t)nnz * sizeof(int32_t), s);
      if (!*indices)
        e = hipErrorOutOfMemory;
    }
    if (e == hipSuccess)
      e = launch_row_prefix_copy(full_off, full_ind, nq, offsets, *indices, s);
    if (e == hipSuccess)
      e = sync_stream(s);
  }
  out_free(full_ind, s);
  dfree(full_off, s);
  dfree(scan_ws, s);
  if (st != AXB_OK)
    return st;
  if (e != hipSuccess)
  {
    out_free(*indices, s);
    *indices = nullptr;
    return hip_fail(e, "axb_bvh_query_nearest_rays");
  }
  st = check_kernel_error(err.p, s, "axb_bvh_query_nearest_rays");
  if (st != AXB_OK)
  {
    out_free(*indices, s);
    *indices = nullptr;
    return st;
  }
  *nnz_host = nnz;
  return AXB_OK;
}

axb_status axb_make_rays(const float *origins, const float *directions, int64_t n,
                         float *rays, axb_stream_t stream_)
{
  if (n < 0 || (n > 0 && (!origins || !directions || !rays)))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_make_rays: NULL argument or n < 0");
  if (n == 0)
    return AXB_OK;
  hipStream_t s = (hipStream_t)stream_;
  uint32_t *err = (uint32_t *)dalloc(sizeof(uint32_t), s);
  if (!err)
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "axb_make_rays: out of device memory");
  hipError_t e = fill_u32(err, 0u, 1, s);
  if (e == hipSuccess)
    e = launch_make_rays(origins, directions, n, rays, err, s);
  uint32_t flags = 0;
  if (e == hipSuccess)
    e = hipMemcpyAsync(&flags, err, sizeof(flags), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = sync_stream(s);
  dfree(err, s);
  if (e != hipSuccess)
    return hip_fail(e, "axb_make_rays");
  if (flags)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_make_rays: a direction is the zero vector");
  return AXB_OK;
}

axb_status axb_bvh_query_intersects_radius(const axb_bvh *t, const float *points, int64_t nq,
                                           float radius, axb_traversal_policy policy,
                                           int64_t *offsets, int32_t **indices,
                                           int64_t *nnz_host, axb_stream_t stream)
{
  // PredicateHelpers.hpp:102: KOKKOS_ASSERT(r > 0)
  if (!(radius > 0.f))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_query_intersects_radius: radius must be > 0");
  return query_spatial(t, PRED_RADIUS, points, nq, radius, policy, offsets, indices, nnz_host,
                       stream);
}

axb_status axb_bvh_view(const axb_bvh *t, axb_tree_view *out)
{
  if (!t || !out)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_view: NULL argument");
  const TreeView tv = view_of(t, nullptr);
  out->slots = t->n > 0 ? (const void *)tv.slots : nullptr;
  out->end_slot = t->n > 0 ? (int64_t)tv.total : 0;
  out->n = t->n;
  out->first_leaf_prefix = t->n > 0 ? tv.C : nullptr;
  for (int d = 0; d < 3; ++d)
  {
    out->scene[d] = tv.scene.lo[d];
    out->scene[3 + d] = tv.scene.hi[d];
    out->q16_scale[d] = tv.ss[d];
  }
  out->geometry = t->geometry;
  out->node_format = t->node_format;
  return AXB_OK;
}

axb_status axb_bvh_predicate_order(const axb_bvh *t, int32_t predicate_kind,
                                   const float *preds, int64_t nq, uint32_t *order,
                                   axb_stream_t stream_)
{
  if (!t || (nq > 0 && (!preds || !order)))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_predicate_order: NULL argument");
  if (predicate_kind != AXB_PREDICATE_SPHERES && predicate_kind != AXB_PREDICATE_BOXES &&
      predicate_kind != PRED_RADIUS && predicate_kind != AXB_PREDICATE_RAYS &&
      predicate_kind != AXB_PREDICATE_ORDERED_RAYS && ordered_geom_kind(predicate_kind) < 0)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_predicate_order: unknown predicate kind");
  // the Morton code of a predicate's centroid: rays by their origin, an
  // ordered sphere / box / point like the unordered one
  if (predicate_kind == AXB_PREDICATE_ORDERED_RAYS)
    predicate_kind = PRED_RAYS;
  else if (predicate_kind == AXB_PREDICATE_ORDERED_SPHERES)
    predicate_kind = PRED_SPHERES;
  else if (predicate_kind == AXB_PREDICATE_ORDERED_BOXES)
    predicate_kind = PRED_BOXES;
  else if (predicate_kind == AXB_PREDICATE_ORDERED_POINTS)
    predicate_kind = PRED_RADIUS;
  // the sort's look-back words carry 30-bit counts
  if (nq < 0 || nq >= MAX_SORTED_PREDICATES)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_predicate_order: nq must be in [0, 2^30)");
  hipStream_t s = (hipStream_t)stream_;
  if (nq == 0)
    return AXB_OK;
  if (nq == 1 || t->n == 0)
  {
    AXB_HIP_TRY(launch_iota_u32(order, nq, s));
    return AXB_OK;
  }
  CallErr err(s);
  if (!err.p)
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "axb_bvh_predicate_order: out of device memory");
  uint32_t *ord = nullptr;
  void *blk = nullptr;
  axb_status st = predicate_order(t, err.p, predicate_kind, preds, nq, s, &ord, &blk, true);
  if (st != AXB_OK)
    return st;
  hipError_t e = hipMemcpyAsync(order, ord, (size_t)nq * sizeof(uint32_t),
                                hipMemcpyDeviceToDevice, s);
  dfree(blk, s);
  if (e != hipSuccess)
    return hip_fail(e, "axb_bvh_predicate_order");
  return check_kernel_error(err.p, s, "axb_bvh_predicate_order");
}

axb_status axb_exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n,
                                  axb_stream_t stream_)
{
  if (n < 0 || (n > 0 && (!in || !out)))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_exclusive_scan_i64: bad argument");
  if (n == 0)
    return AXB_OK;
  hipStream_t s = (hipStream_t)stream_;
  void *ws = dalloc(scan_workspace_bytes(n), s);
  uint32_t *err = (uint32_t *)dalloc(sizeof(uint32_t), s);
  hipError_t e = (ws && err) ? hipSuccess : hipErrorOutOfMemory;
  if (e == hipSuccess)
    e = fill_u32(err, 0u, 1, s);
  if (e == hipSuccess)
    e = exclusive_scan_i64(in, out, n, ws, err, s);
  axb_status st = AXB_OK;
  if (e == hipSuccess)
    st = check_kernel_error(err, s, "axb_exclusive_scan_i64");
  dfree(ws, s);
  dfree(err, s);
  if (e != hipSuccess)
    return hip_fail(e, "axb_exclusive_scan_i64");
  return st;
}

axb_status axb_bvh_neighbor_list(const axb_bvh *t, float radius, int32_t full,
                                 axb_traversal_policy policy, int64_t *offsets,
                                 int32_t **indices, int64_t *nnz_host, axb_stream_t stream_)
{
  Region region(full ? "ArborX::Experimental::FullNeighborList"
                     : "ArborX::Experimental::HalfNeighborList");
  if (!t || !offsets || !indices || !nnz_host)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_neighbor_list: NULL argument");
  if (t->geometry != AXB_GEOMETRY_POINTS)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_neighbor_list: the tree must hold points");
  if (!(radius >= 0.f))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_neighbor_list: radius must be >= 0");
  hipStream_t s = (hipStream_t)stream_;
  *indices = nullptr;
  *nnz_host = 0;
  const int64_t n = t->n;
  if (n <= 1) // HalfTraversal.hpp:33-44: no pairs
  {
    AXB_HIP_TRY(fill_u64((uint64_t *)offsets, 0ull, n + 1, s));
    *indices = (int32_t *)out_alloc(0, s);
    AXB_HIP_TRY(sync_stream(s));
    return AXB_OK;
  }
  CallErr err(s);
  if (!err.p)
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "axb_bvh_neighbor_list: out of device memory");
  const TreeView tv = view_of(t, err.p);
  uint32_t *rowof = (uint32_t *)dalloc((size_t)n * sizeof(uint32_t), s);
  if (!rowof)
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "axb_bvh_neighbor_list: out of device memory");
  const bool f = full != 0;
  // lanes walk the leaves in tree order (already Morton-sorted)
  auto first = [&](int mode, int32_t *out, int32_t B) {
    return launch_nlist(tv, t->node_format, f, radius, n, nullptr, offsets, out, B, rowof, mode,
                        s);
  };
  // the overflow list holds leaf positions
  auto fill = [&](const uint32_t *list, int64_t n_list, int32_t *out) {
    return launch_nlist(tv, t->node_format, f, radius, n_list, list, offsets, out, 0, nullptr,
                        MODE_FILL, s);
  };
  // rows of 96, as for queries: the 2M-atom FCC lattice's ~83 neighbours
  // per row take 5.1 ms with rows of 64 (most rows walked twice), 2.4 ms
  // with 96; ~32 uniform neighbours 6.2 vs 6.3 ms at 10M points
  axb_status st = crs_query(err.p, n, policy, rowof, true, first, fill, offsets, indices,
                            nnz_host, s, "axb_bvh_neighbor_list", nullptr, 96);
  dfree(rowof, s);
  return st;
}

axb_status axb_find_neighbor_list(const float *points, int64_t n, float radius, int32_t full,
                                  int64_t *offsets, int32_t **indices, int64_t *nnz_host,
                                  axb_stream_t stream)
{
  axb_bvh *t = nullptr;
  axb_status st = axb_bvh_create(points, n, AXB_GEOMETRY_POINTS, AXB_NODES_Q16, 0, stream, &t);
  if (st == AXB_ERR_INVALID_ARGUMENT && n > 1 && points)
    // scene box without finite extent: 32-byte nodes have no such limit
    st = axb_bvh_create(points, n, AXB_GEOMETRY_POINTS, AXB_NODES_FP32, 0, stream, &t);
  if (st != AXB_OK)
    return st;
  st = axb_bvh_neighbor_list(t, radius, full, axb_traversal_policy{0, 1}, offsets, indices,
                             nnz_host, stream);
  axb_bvh_destroy(t);
  return st;
}

axb_status axb_expand_half_to_full(const int64_t *offsets, const int32_t *indices, int64_t n,
                                   int64_t *out_offsets, int32_t **out_indices,
                                   int64_t *nnz_host, axb_stream_t stream_)
{
  Region region("ArborX::Experimental::HalfToFull");
  if (!offsets || !out_offsets || !out_indices || !nnz_host)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_expand_half_to_full: NULL argument");
  if (n < 0 || n >= (int64_t(1) << 31))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_expand_half_to_full: n must be in [0, 2^31)");
  hipStream_t s = (hipStream_t)stream_;
  *out_indices = nullptr;
  *nnz_host = 0;
  int64_t half = 0;
  AXB_HIP_TRY(hipMemcpyAsync(&half, offsets + n, sizeof(half), hipMemcpyDeviceToHost, s));
  AXB_HIP_TRY(sync_stream(s));
  if (half > 0 && !indices)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_expand_half_to_full: indices is NULL");
  int64_t *cursor = (int64_t *)dalloc((size_t)(n + 1) * sizeof(int64_t), s);
  void *scan_ws = dalloc(scan_workspace_bytes(n + 1), s);
  uint32_t *err = (uint32_t *)dalloc(sizeof(uint32_t), s);
  hipError_t e = (cursor && scan_ws && err) ? hipSuccess : hipErrorOutOfMemory;
  if (e == hipSuccess)
    e = fill_u32(err, 0u, 1, s);
  if (e == hipSuccess)
    e = fill_u64((uint64_t *)out_offsets, 0ull, n + 1, s);
  if (e == hipSuccess)
    e = launch_h2f_count(offsets, indices, n, out_offsets, err, s);
  if (e == hipSuccess)
    e = exclusive_scan_i64(out_offsets, out_offsets, n + 1, scan_ws, err, s);
  int64_t m = 0;
  if (e == hipSuccess)
    e = hipMemcpyAsync(&m, out_offsets + n, sizeof(m), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = sync_stream(s);
  if (e == hipSuccess)
  {
    *out_indices = (int32_t *)out_alloc((size_t)m * sizeof(int32_t), s);
    if (!*out_indices)
      e = hipErrorOutOfMemory;
  }
  if (e == hipSuccess && m > 0)
    e = launch_h2f_fill(offsets, indices, n, out_offsets, cursor, *out_indices, s);
  if (e == hipSuccess)
    e = sync_stream(s);
  axb_status st = AXB_OK;
  if (e == hipSuccess)
    st = check_kernel_error(err, s, "axb_expand_half_to_full");
  dfree(cursor, s);
  dfree(scan_ws, s);
  dfree(err, s);
  if (e != hipSuccess || st != AXB_OK)
  {
    out_free(*out_indices, s);
    *out_indices = nullptr;
    return e != hipSuccess ? hip_fail(e, "axb_expand_half_to_full") : st;
  }
  *nnz_host = m;
  return AXB_OK;
}

axb_status axb_bvh_query_nearest(const axb_bvh *t, const float *points, int64_t nq, int32_t k,
                                 axb_traversal_policy policy, int64_t *offsets,
                                 int32_t **indices, float **distances, int64_t *nnz_host,
                                 axb_stream_t stream_)
{
  Region region("ArborX::BVH::query_crs");
  Region tag_region("ArborX::BVH::query::nearest");
  if (!t || !offsets || !indices || !nnz_host)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_query_nearest: NULL argument");
  if (nq < 0 || nq >= (int64_t(1) << 31))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_query_nearest: nq must be in [0, 2^31)");
  if (nq > 0 && !points)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_query_nearest: points is NULL");
  if (policy.sort_predicates && nq >= MAX_SORTED_PREDICATES)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT,
             "axb_bvh_query_nearest: sort_predicates needs fewer than 2^30 predicates");
  hipStream_t s = (hipStream_t)stream_;
  const int64_t per = (k < 1 || t->n == 0) ? 0 : std::min<int64_t>(k, t->n);
  const int64_t nnz = nq * per;
  if (nnz >= (int64_t(1) << 40))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bvh_query_nearest: result too large");
  *indices = nullptr;
  if (distances)
    *distances = nullptr;
  CallErr err(s);
  if (!err.p)
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "axb_bvh_query_nearest: out of device memory");
  AXB_HIP_TRY(launch_knn_offsets(offsets, nq, per, s));
  *indices = (int32_t *)out_alloc((size_t)nnz * sizeof(int32_t), s);
  float *dist = distances ? (float *)out_alloc((size_t)nnz * sizeof(float), s) : nullptr;
  float *tmp_dist = nullptr;
  if (!*indices || (distances && !dist))
  {
    out_free(*indices, s);
    out_free(dist, s);
    *indices = nullptr;
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "axb_bvh_query_nearest: out of device memory");
  }
  hipError_t e = hipSuccess;
  uint32_t *order = nullptr;
  void *order_blk = nullptr;
  if (nnz > 0)
  {
    if (k > 64 && !dist)
    {
      tmp_dist = (float *)dalloc((size_t)nnz * sizeof(float), s);
      if (!tmp_dist)
        e = hipErrorOutOfMemory;
    }
    if (e == hipSuccess && policy.sort_predicates)
    {
      Region perm_region("ArborX::BVH::query::nearest::compute_permutation");
      axb_status st = predicate_order(t, err.p, PRED_RADIUS, points, nq, s, &order, &order_blk);
      if (st != AXB_OK)
      {
        out_free(*indices, s);
        out_free(dist, s);
        dfree(tmp_dist, s);
        *indices = nullptr;
        return st;
      }
    }
    if (e == hipSuccess)
      e = launch_knn(view_of(t, err.p), t->node_format, t->geometry, points, nq, k, order, *indices,
                     dist ? dist : tmp_dist, s);
  }
  if (e == hipSuccess)
    e = sync_stream(s);
  dfree(order_blk, s);
  dfree(tmp_dist, s);
  axb_status st = e == hipSuccess ? check_kernel_error(err.p, s, "axb_bvh_query_nearest")
                                  : hip_fail(e, "axb_bvh_query_nearest");
  if (st != AXB_OK)
  {
    out_free(*indices, s);
    out_free(dist, s);
    *indices = nullptr;
    return st;
  }
  if (distances)
    *distances = dist;
  *nnz_host = nnz;
  return AXB_OK;
}

axb_status axb_profile_enable(int32_t flags)
{
  axb_profile_clear();
  if ((flags & 2) && !g_visits)
  {
    hipError_t e = hipMalloc((void **)&g_visits, 12 * sizeof(unsigned long long));
    if (e != hipSuccess)
      return hip_fail(e, "axb_profile_enable");
  }
  if (g_visits)
  {
    hipError_t e = hipMemset(g_visits, 0, 12 * sizeof(unsigned long long));
    if (e == hipSuccess)
      e = hipDeviceSynchronize();
    if (e != hipSuccess)
      return hip_fail(e, "axb_profile_enable");
  }
  g_prof_timing.store((flags & 1) != 0);
  g_prof_visits.store((flags & 2) != 0);
  return AXB_OK;
}

axb_status axb_profile_clear(void)
{
  std::lock_guard<std::mutex> lock(g_prof_mutex);
  for (auto &r : g_prof_recs)
  {
    if (r.host_ms >= 0.0)
      continue;
    (void)hipEventSynchronize(r.b);
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  g_prof_recs.clear();
  return AXB_OK;
}

int32_t axb_profile_num_records(void)
{
  std::lock_guard<std::mutex> lock(g_prof_mutex);
  return (int32_t)g_prof_recs.size();
}

axb_status axb_profile_record(int32_t i, char *name, int32_t name_len, double *ms)
{
  std::lock_guard<std::mutex> lock(g_prof_mutex);
  if (i < 0 || i >= (int32_t)g_prof_recs.size() || !ms)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_profile_record: bad index");
  const ProfRec &r = g_prof_recs[i];
  if (r.host_ms >= 0.0)
    *ms = r.host_ms;
  else
  {
    hipError_t e = hipEventSynchronize(r.b);
    float t = 0.f;
    if (e == hipSuccess)
      e = hipEventElapsedTime(&t, r.a, r.b);
    if (e != hipSuccess)
      return hip_fail(e, "axb_profile_record");
    *ms = t;
  }
  if (name && name_len > 0)
  {
    std::strncpy(name, r.name.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  return AXB_OK;
}

axb_status axb_profile_visits(int64_t out_host[12])
{
  if (!out_host)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_profile_visits: NULL");
  for (int i = 0; i < 12; ++i)
    out_host[i] = 0;
  if (!g_visits)
    return AXB_OK;
  unsigned long long h[12];
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess)
    e = hipMemcpy(h, g_visits, sizeof(h), hipMemcpyDeviceToHost);
  if (e != hipSuccess)
    return hip_fail(e, "axb_profile_visits");
  for (int i = 0; i < 12; ++i)
    out_host[i] = (int64_t)h[i];
  return AXB_OK;
}

axb_status axb_pair_energy(const float *xyz, int64_t n, const int64_t *offsets,
                           const int32_t *indices, float cutoff, double *energy, float *grad,
                           axb_stream_t stream)
{
  if (n < 0 || (n > 0 && (!xyz || !offsets)))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_pair_energy: bad arguments");
  if (!(cutoff > 0.f))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_pair_energy: cutoff must be > 0");
  hipStream_t s = (hipStream_t)stream;
  // Rows in Morton order of the points (scene box of xyz, 32-bit codes):
  // the neighbours gathered by a wavefront then overlap, whatever the order
  // of the caller's points.  Worth its sort only when the coordinates do
  // not stay in the 256 MB Infinity Cache anyway (measured: 100M random-
  // order points 100 -> 78 ms per pass; 2M lattice points +0.4 ms).
  uint32_t *order = nullptr;
  void *order_blk = nullptr;
  float *scene = nullptr;
  uint32_t *d_err = nullptr;
  hipError_t e = hipSuccess;
  if (12 * n > (int64_t(64) << 20) && n < MAX_SORTED_PREDICATES)
  {
    scene = (float *)dalloc(6 * sizeof(float), s);
    d_err = (uint32_t *)dalloc(sizeof(uint32_t), s);
    e = (scene && d_err) ? hipSuccess : hipErrorOutOfMemory;
    if (e == hipSuccess)
      e = fill_u32(d_err, 0u, 1, s);
    if (e == hipSuccess)
      e = scene_box(xyz, n, AXB_GEOMETRY_POINTS, scene, s);
    if (e == hipSuccess &&
        morton_order(scene, d_err, PRED_RADIUS, xyz, n, s, &order, &order_blk) != AXB_OK)
      e = hipErrorOutOfMemory;
  }
  if (e == hipSuccess)
    e = pair_energy(xyz, n, offsets, indices, cutoff, order, energy, grad, s);
  axb_status st = e == hipSuccess ? AXB_OK : hip_fail(e, "axb_pair_energy");
  if (st == AXB_OK && d_err)
    st = check_kernel_error(d_err, s, "axb_pair_energy");
  dfree(order_blk, s);
  dfree(scene, s);
  dfree(d_err, s);
  return st;
}

axb_status axb_pair_energy_noalloc(const float *xyz, int64_t n, const int64_t *offsets,
                                   const int32_t *indices, float cutoff, double *energy,
                                   float *grad, axb_stream_t stream)
{
  static_assert(AXB_PAIR_ENERGY_NOALLOC_MAX_N == ENERGY_SCRATCH_BLOCKS * 256,
                "scratch size and the documented limit differ");
  if (n < 0 || (n > 0 && (!xyz || !offsets)))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_pair_energy_noalloc: bad arguments");
  if (n > AXB_PAIR_ENERGY_NOALLOC_MAX_N)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT,
             "axb_pair_energy_noalloc: n exceeds AXB_PAIR_ENERGY_NOALLOC_MAX_N (2^27 points)");
  if (!(cutoff > 0.f))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_pair_energy_noalloc: cutoff must be > 0");
  const hipError_t e = pair_energy_noalloc(xyz, n, offsets, indices, cutoff, energy, grad,
                                           (hipStream_t)stream);
  if (e != hipSuccess)
    return hip_fail(e, "axb_pair_energy_noalloc");
  return AXB_OK;
}

axb_status axb_pair_energy_fd(const float *xyz, int64_t n, const int64_t *offsets,
                              const int32_t *indices, float cutoff, const int32_t *atoms,
                              int64_t m, double h, double *fd, axb_stream_t stream)
{
  if (n < 0 || m < 0 || (m > 0 && (!xyz || !offsets || !indices || !atoms || !fd)))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_pair_energy_fd: bad arguments");
  if (!(h > 0.0) || !(cutoff > 0.f))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_pair_energy_fd: h and cutoff must be > 0");
  hipError_t e =
      pair_energy_fd(xyz, offsets, indices, cutoff, atoms, m, h, fd, (hipStream_t)stream);
  if (e != hipSuccess)
    return hip_fail(e, "axb_pair_energy_fd");
  return AXB_OK;
}


// ------------------------------------------------------------ brute force
// ArborX::BruteForce (src/spatial/ArborX_BruteForce.hpp): the values (a
// device copy), their bounds and the kernels' error word.
struct axb_bruteforce
{
  int64_t n = 0;
  int geometry = 0;
  float *values = nullptr;
  float bounds[6];
  hipStream_t stream = nullptr;
  uint32_t *d_err = nullptr;
};

axb_status axb_bruteforce_create(const float *values, int64_t n, int32_t geometry,
                                 axb_stream_t stream_, axb_bruteforce **out)
{
  Region region("ArborX::BruteForce::BruteForce");
  if (!out)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bruteforce_create: out is NULL");
  *out = nullptr;
  if (n < 0 || n >= (int64_t(1) << 31))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bruteforce_create: n must be in [0, 2^31)");
  if (geometry != AXB_GEOMETRY_POINTS && geometry != AXB_GEOMETRY_BOXES)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bruteforce_create: unknown geometry");
  if (n > 0 && !values)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bruteforce_create: values is NULL");
  hipStream_t s = (hipStream_t)stream_;
  const size_t bytes = (size_t)n * (geometry == AXB_GEOMETRY_POINTS ? 3 : 6) * sizeof(float);
  axb_bruteforce *bf = new axb_bruteforce;
  bf->n = n;
  bf->geometry = geometry;
  bf->stream = s;
  bf->d_err = (uint32_t *)dalloc(sizeof(uint32_t), s);
  bf->values = n > 0 ? (float *)dalloc(bytes, s) : nullptr;
  float *scene = (float *)dalloc(6 * sizeof(float), s);
  hipError_t e = (bf->d_err && scene && (n == 0 || bf->values)) ? hipSuccess : hipErrorOutOfMemory;
  if (e == hipSuccess)
    e = fill_u32(bf->d_err, 0u, 1, s);
  if (e == hipSuccess && n > 0)
    e = hipMemcpyAsync(bf->values, values, bytes, hipMemcpyDeviceToDevice, s);
  // bounds: the reduction over the values (BruteForceImpl.hpp:34-50)
  if (e == hipSuccess)
    e = scene_box(bf->values, n, geometry, scene, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(bf->bounds, scene, sizeof(bf->bounds), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = sync_stream(s);
  dfree(scene, s);
  if (e != hipSuccess)
  {
    axb_status st = hip_fail(e, "axb_bruteforce_create");
    axb_bruteforce_destroy(bf);
    return st;
  }
  *out = bf;
  return AXB_OK;
}

void axb_bruteforce_destroy(axb_bruteforce *bf)
{
  if (!bf)
    return;
  dfree(bf->values, bf->stream);
  dfree(bf->d_err, bf->stream);
  delete bf;
}

int64_t axb_bruteforce_size(const axb_bruteforce *bf) { return bf ? bf->n : 0; }

axb_status axb_bruteforce_bounds(const axb_bruteforce *bf, float out_host[6])
{
  if (!bf || !out_host)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bruteforce_bounds: NULL argument");
  std::memcpy(out_host, bf->bounds, sizeof(bf->bounds));
  return AXB_OK;
}

axb_status axb_bruteforce_view(const axb_bruteforce *bf, axb_values_view *out)
{
  if (!bf || !out)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bruteforce_view: NULL argument");
  out->values = bf->values;
  out->n = bf->n;
  out->geometry = bf->geometry;
  return AXB_OK;
}

static axb_status bf_query_spatial(const axb_bruteforce *bf, int pred_kind, const float *preds,
                                   int64_t nq, float radius, axb_traversal_policy policy,
                                   int64_t *offsets, int32_t **indices, int64_t *nnz_host,
                                   axb_stream_t stream_)
{
  Region region("ArborX::BruteForce::query_crs");
  Region tag_region("ArborX::BruteForce::query::spatial");
  if (!bf || !offsets || !indices || !nnz_host)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "bruteforce query: NULL argument");
  if (nq < 0 || nq >= (int64_t(1) << 31))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "bruteforce query: nq must be in [0, 2^31)");
  if (nq > 0 && !preds)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "bruteforce query: predicates is NULL");
  hipStream_t s = (hipStream_t)stream_;
  *indices = nullptr;
  *nnz_host = 0;
  if (nq == 0 || bf->n == 0)
  {
    AXB_HIP_TRY(fill_u64((uint64_t *)offsets, 0ull, nq + 1, s));
    *indices = (int32_t *)out_alloc(0, s);
    AXB_HIP_TRY(sync_stream(s));
    return AXB_OK;
  }
  CallErr err(s);
  if (!err.p)
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "bruteforce query: out of device memory");
  // predicates in the caller's order: the values are read in the same order
  // by every lane, so there is no locality to gain from sorting them
  auto first = [&](int mode, int32_t *out, int32_t B) {
    return launch_bf_spatial(bf->values, bf->n, bf->geometry, pred_kind, preds, nq, radius,
                             nullptr, offsets, out, B, mode, err.p, s);
  };
  auto fill = [&](const uint32_t *list, int64_t n_list, int32_t *out) {
    return launch_bf_spatial(bf->values, bf->n, bf->geometry, pred_kind, preds, n_list, radius,
                             list, offsets, out, 0, MODE_FILL, err.p, s);
  };
  return crs_query(err.p, nq, policy, nullptr, false, first, fill, offsets, indices,
                   nnz_host, s, "bruteforce query");
}

axb_status axb_bruteforce_query_intersects(const axb_bruteforce *bf, int32_t predicate_kind,
                                          const float *predicates, int64_t nq,
                                          axb_traversal_policy policy, int64_t *offsets,
                                          int32_t **indices, int64_t *nnz_host,
                                          axb_stream_t stream)
{
  if (predicate_kind != AXB_PREDICATE_SPHERES && predicate_kind != AXB_PREDICATE_BOXES &&
      predicate_kind != AXB_PREDICATE_RAYS)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT,
             "axb_bruteforce_query_intersects: unknown predicate kind");
  if (predicate_kind == AXB_PREDICATE_RAYS && bf && bf->geometry != GEOM_BOXES && bf->n > 0)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT,
             "axb_bruteforce_query_intersects: ray predicates need boxes");
  return bf_query_spatial(bf, predicate_kind, predicates, nq, 0.f, policy, offsets, indices,
                          nnz_host, stream);
}

axb_status axb_bruteforce_query_intersects_radius(const axb_bruteforce *bf, const float *points,
                                                 int64_t nq, float radius,
                                                 axb_traversal_policy policy, int64_t *offsets,
                                                 int32_t **indices, int64_t *nnz_host,
                                                 axb_stream_t stream)
{
  // PredicateHelpers.hpp:102: KOKKOS_ASSERT(r > 0)
  if (!(radius > 0.f))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT,
             "axb_bruteforce_query_intersects_radius: radius must be > 0");
  return bf_query_spatial(bf, PRED_RADIUS, points, nq, radius, policy, offsets, indices,
                          nnz_host, stream);
}

axb_status axb_bruteforce_query_nearest_rays(const axb_bruteforce *bf, const float *rays,
                                             int64_t nq, int32_t k, axb_traversal_policy policy,
                                             int64_t *offsets, int32_t **indices,
                                             int64_t *nnz_host, axb_stream_t stream_)
{
  Region region("ArborX::BruteForce::query_crs");
  Region tag_region("ArborX::BruteForce::query::nearest");
  if (!bf || !offsets || !indices || !nnz_host)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bruteforce_query_nearest_rays: NULL argument");
  if (nq < 0 || nq >= (int64_t(1) << 31))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT,
             "axb_bruteforce_query_nearest_rays: nq must be in [0, 2^31)");
  if (nq > 0 && !rays)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bruteforce_query_nearest_rays: rays is NULL");
  if (bf->geometry != GEOM_BOXES && bf->n > 0)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT,
             "axb_bruteforce_query_nearest_rays: ray predicates need boxes");
  hipStream_t s = (hipStream_t)stream_;
  *indices = nullptr;
  *nnz_host = 0;
  const int kk = k > 0 ? k : 0; // TreeTraversal.hpp:143-178: k < 1 returns nothing
  if (nq == 0 || bf->n == 0 || kk == 0)
  {
    AXB_HIP_TRY(fill_u64((uint64_t *)offsets, 0ull, nq + 1, s));
    *indices = (int32_t *)out_alloc(0, s);
    AXB_HIP_TRY(sync_stream(s));
    return AXB_OK;
  }
  CallErr err(s);
  if (!err.p)
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "bruteforce nearest rays: out of device memory");
  auto first = [&](int mode, int32_t *out, int32_t B) {
    return launch_bf_ray_nearest(bf->values, bf->n, rays, nq, kk, nullptr, offsets, out, B,
                                 mode, err.p, s);
  };
  auto fill = [&](const uint32_t *list, int64_t n_list, int32_t *out) {
    return launch_bf_ray_nearest(bf->values, bf->n, rays, n_list, kk, list, offsets, out, 0,
                                 MODE_FILL, err.p, s);
  };
  return crs_query(err.p, nq, policy, nullptr, false, first, fill, offsets, indices,
                   nnz_host, s, "bruteforce nearest rays");
}

axb_status axb_bruteforce_query_nearest(const axb_bruteforce *bf, const float *points,
                                       int64_t nq, int32_t k, axb_traversal_policy policy,
                                       int64_t *offsets, int32_t **indices, float **distances,
                                       int64_t *nnz_host, axb_stream_t stream_)
{
  Region region("ArborX::BruteForce::query_crs");
  Region tag_region("ArborX::BruteForce::query::nearest");
  (void)policy; // no traversal order to choose: every value is visited
  if (!bf || !offsets || !indices || !nnz_host)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bruteforce_query_nearest: NULL argument");
  if (nq < 0 || nq >= (int64_t(1) << 31))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bruteforce_query_nearest: nq must be in [0, 2^31)");
  if (nq > 0 && !points)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bruteforce_query_nearest: points is NULL");
  hipStream_t s = (hipStream_t)stream_;
  const int64_t per = (k < 1 || bf->n == 0) ? 0 : std::min<int64_t>(k, bf->n);
  const int64_t nnz = nq * per;
  if (nnz >= (int64_t(1) << 40))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_bruteforce_query_nearest: result too large");
  *indices = nullptr;
  if (distances)
    *distances = nullptr;
  CallErr err(s);
  if (!err.p)
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "axb_bruteforce_query_nearest: out of device memory");
  AXB_HIP_TRY(launch_knn_offsets(offsets, nq, per, s));
  *indices = (int32_t *)out_alloc((size_t)nnz * sizeof(int32_t), s);
  float *dist = distances ? (float *)out_alloc((size_t)nnz * sizeof(float), s) : nullptr;
  if (!*indices || (distances && !dist))
  {
    out_free(*indices, s);
    out_free(dist, s);
    *indices = nullptr;
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "axb_bruteforce_query_nearest: out of device memory");
  }
  hipError_t e = hipSuccess;
  float *tmp_dist = nullptr;
  if (nnz > 0)
  {
    if (k > 64 && !dist) // the heap keeps its distances in the output row
    {
      tmp_dist = (float *)dalloc((size_t)nnz * sizeof(float), s);
      if (!tmp_dist)
        e = hipErrorOutOfMemory;
    }
    if (e == hipSuccess)
      e = launch_bf_nearest(bf->values, bf->n, bf->geometry, points, nq, k, *indices,
                            dist ? dist : tmp_dist, s);
  }
  if (e == hipSuccess)
    e = sync_stream(s);
  dfree(tmp_dist, s);
  axb_status st = e == hipSuccess ? check_kernel_error(err.p, s, "axb_bruteforce_query_nearest")
                                  : hip_fail(e, "axb_bruteforce_query_nearest");
  if (st != AXB_OK)
  {
    out_free(*indices, s);
    out_free(dist, s);
    *indices = nullptr;
    return st;
  }
  if (distances)
    *distances = dist;
  *nnz_host = nnz;
  return AXB_OK;
}

} // extern "C"

namespace
{
// Record j of out = {values[indices[j]] (VF floats), indices[j]}; an index
// outside [0, n_values) sets *bad and leaves its record unwritten.
template <int VF>
__global__ __launch_bounds__(256) void k_gather_pairs(const float *__restrict__ values,
                                                      int64_t n_values,
                                                      const int32_t *__restrict__ indices,
                                                      int64_t nnz, uint32_t *__restrict__ out,
                                                      uint32_t *__restrict__ bad)
{
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nnz)
    return;
  const int32_t i = indices[j];
  if (i < 0 || i >= n_values)
  {
    *bad = 1u;
    return;
  }
  uint32_t *o = out + j * (VF + 1);
#pragma unroll
  for (int d = 0; d < VF; ++d)
    o[d] = __float_as_uint(values[(int64_t)i * VF + d]);
  o[VF] = (uint32_t)i;
}
} // namespace

extern "C" {

axb_status axb_gather_pairs(const float *values, int32_t value_floats, int64_t n_values,
                            const int32_t *indices, int64_t nnz, void *out,
                            axb_stream_t stream_)
{
  if (value_floats != 3 && value_floats != 6)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT,
             "axb_gather_pairs: values are points (3 floats) or boxes (6 floats)");
  if (nnz < 0 || n_values < 0 || (nnz > 0 && (!values || !indices || !out)))
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_gather_pairs: bad argument");
  if (nnz == 0)
    return AXB_OK;
  hipStream_t s = (hipStream_t)stream_;
  uint32_t *bad = (uint32_t *)dalloc(sizeof(uint32_t), s);
  if (!bad)
    AXB_FAIL(AXB_ERR_OUT_OF_MEMORY, "axb_gather_pairs: out of device memory");
  uint32_t bad_host = 0;
  hipError_t e = fill_u32(bad, 0u, 1, s);
  const unsigned grid = (unsigned)div_up(nnz, 256);
  if (e == hipSuccess && value_floats == 3)
  {
    AXB_LAUNCH(k_gather_pairs<3>, dim3(grid), dim3(256), 0, s, values, n_values, indices, nnz,
               (uint32_t *)out, bad);
    e = hipGetLastError();
  }
  else if (e == hipSuccess)
  {
    AXB_LAUNCH(k_gather_pairs<6>, dim3(grid), dim3(256), 0, s, values, n_values, indices, nnz,
               (uint32_t *)out, bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(&bad_host, bad, sizeof(bad_host), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = sync_stream(s);
  dfree(bad, s);
  if (e != hipSuccess)
    return hip_fail(e, "axb_gather_pairs");
  if (bad_host)
    AXB_FAIL(AXB_ERR_INVALID_ARGUMENT, "axb_gather_pairs: an index is outside [0, n_values)");
  return AXB_OK;
}

} // extern "C"
