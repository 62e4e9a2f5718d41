// axb_internal.hpp -- declarations shared by the library's translation units.
#pragma once

#include "arborx_mi355x_device.hpp"

#include <chrono>

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace axb
{

// ------------------------------------------------------------ errors
void set_error(const std::string &msg);
#define AXB_HIP_TRY(expr)                                                      \
  do                                                                           \
  {                                                                            \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
    {                                                                          \
      ::axb::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));    \
      return AXB_ERR_HIP;                                                      \
    }                                                                          \
  } while (0)

// Stream-ordered device allocation (hipMallocAsync on `stream`).
void *dalloc(size_t bytes, hipStream_t stream);
void dfree(void *p, hipStream_t stream);

inline int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ------------------------------------------------------------ profiling
// While per-kernel timing is enabled (axb_profile_enable), every launch is
// bracketed by HIP events recorded on its own stream.
struct ProfScope
{
  ProfScope(const char *name, hipStream_t s);
  ~ProfScope();
  const char *name_;
  hipStream_t s_;
  bool on_;
  hipEvent_t a_, b_;
};
// Named range on the host's timeline, around a public entry point or one of
// its phases, under the reference's profiling region labels (Kokkos regions,
// e.g. ArborX_LinearBVH.hpp:201-256, ArborX_CrsGraphWrapperImpl.hpp:391-410):
// roctx ranges, which `rocprofv3 --marker-trace` records and which cost one
// call each when no tool listens.  next() ends the range and opens the
// following phase's.
struct Region
{
  explicit Region(const char *label);
  ~Region();
  void next(const char *label);
  Region(const Region &) = delete;
  Region &operator=(const Region &) = delete;
};
// Host-side interval recorded with the kernel timings (name "host:...").
struct HostScope
{
  explicit HostScope(const char *name);
  ~HostScope();
  const char *name_;
  bool on_;
  std::chrono::steady_clock::time_point t0_;
};
// Scratch buffer of at least `bytes` for one call on stream s.  Each
// (device, stream, slot) keeps one grow-only block between calls (released
// by axb_trim_memory); a call takes it for its duration, and a call that
// finds it taken (another host thread on the same stream) gets a block of
// its own, freed when the lease ends.  p is nullptr when out of memory.
enum WsSlot
{
  WS_ROWS = 0,     // scratch rows of the radius CRS
  WS_KNN_CAND = 1, // kNN candidate rows
  WS_NUM = 2
};
struct WsLease
{
  void *p = nullptr;
  hipStream_t s = nullptr;
  int dev = 0, slot = 0;
  bool cached = false;
  WsLease() = default;
  WsLease(const WsLease &) = delete;
  WsLease &operator=(const WsLease &) = delete;
  // takes a block of at least `bytes` (a lease held already is released first)
  void acquire(hipStream_t s, int slot, size_t bytes);
  void release();
  ~WsLease() { release(); }
};

// Limits that are normally derived from the device, lowered by tests
// (axb_debug_set_limits) to run the multi-launch and fallback paths on
// small inputs.  0 = default.
struct Limits
{
  int64_t knn_batch_max;      // kNN queries per collect launch
  int64_t scratch_rows_bytes; // scratch rows of a buffered CRS query
  int32_t knn_rounds;         // collect rounds before the exact fallback
};
Limits debug_limits();
// hipStreamSynchronize, timed as "host:stream_sync" while profiling.
hipError_t sync_stream(hipStream_t s);
#define AXB_LAUNCH(kernel, grid, block, shmem, stream, ...)                    \
  do                                                                           \
  {                                                                            \
    ::axb::ProfScope axb_ps_(#kernel, stream);                                 \
    hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);       \
  } while (0)
// Device counters {queries, internal-node records read, leaf records read,
// results} while visit counting is enabled, else null.
unsigned long long *visit_counters();

// ------------------------------------------------------------ sort / scan
// The onesweep passes carry 30-bit counts in their look-back words, so a
// sort takes fewer than 2^30 keys (trees: n < 2^30; predicate batches that
// are sorted by Morton code: nq < 2^30).
constexpr int64_t MAX_SORTED_PREDICATES = int64_t(1) << 30;
constexpr int SORT_BLOCK = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_BLOCK * SORT_ITEMS;
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

// Workspace bytes for radix_sort_pairs over n keys with `passes` 8-bit digits.
size_t sort_workspace_bytes(int64_t n, int passes, int key_bytes);

// Stable LSD radix sort of keys (u32 or u64) carrying the permutation of
// original indices.  keys[0] holds the input; after the call keys[0] /
// vals[0] hold the sorted keys and permutation (passes must be even).
// `hist` (passes x 256 uint32 bucket counts) must be filled beforehand by
// the caller's key-generation kernel, or be null to compute it here.
hipError_t radix_sort_pairs_u64(uint64_t *keys0, uint64_t *keys1,
                                uint32_t *vals0, uint32_t *vals1, int64_t n,
                                int passes, uint32_t *hist, void *workspace,
                                uint32_t *err, hipStream_t stream);
hipError_t radix_sort_pairs_u32(uint32_t *keys0, uint32_t *keys1,
                                uint32_t *vals0, uint32_t *vals1, int64_t n,
                                int passes, uint32_t *hist, void *workspace,
                                uint32_t *err, hipStream_t stream);

// Device-wide scans, reduce then scan (in place allowed).  over_flags (when
// given, (n + 7) / 8 bytes): bit q % 8 of byte q / 8 is set when input q
// exceeds over_limit.
size_t scan_workspace_bytes(int64_t n);
hipError_t exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, void *workspace,
                              uint32_t *err, hipStream_t stream, uint8_t *over_flags = nullptr,
                              int64_t over_limit = 0);
hipError_t inclusive_scan_i32(const int32_t *in, int32_t *out, int64_t n,
                              void *workspace, uint32_t *err,
                              hipStream_t stream);

// Fill kernels (used instead of hipMemsetAsync so that the code stays
// graph-capture friendly).
hipError_t fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t stream);
hipError_t launch_iota_u32(uint32_t *p, int64_t n, hipStream_t stream); // p[i] = i
hipError_t fill_u64(uint64_t *p, uint64_t v, int64_t n, hipStream_t stream);


// ------------------------------------------------------------ tree data
// Internal node as produced by the climb (48 B).
struct alignas(16) INode
{
  float lo[3];
  int32_t left_child; // reference encoding: leaf j -> j, internal k -> k + n
  float hi[3];
  int32_t rope; // reference encoding, -1 sentinel
  int32_t first, last, h, pad;
};

struct BuildOut
{
  uint4 *slots;        // max_slots + SLOT_PAD (zeroed padding after max_slots)
  int64_t max_slots;   // every internal node stored
  int64_t total_slots; // records actually stored (SL * n + SI * C[n - 1])
  float *scene;        // 6 floats (device)
  int32_t *C;          // n: inclusive prefix of left-spine counts
  // reference-form topology (kept on request)
  uint64_t *keys;
  uint32_t *perm;
  int32_t *leaf_rope;
  INode *inodes;
};

struct TreeView
{
  const uint4 *slots;
  uint32_t total;
  int32_t n;
  const int32_t *C;
  Box6 scene;
  float ss[3];   // Q16 scale of the scene frame
  uint32_t *err; // kernel error word
  unsigned long long *visits; // optional visit counters (see visit_counters)
};

constexpr int MODE_COUNT = 0;
constexpr int MODE_FILL = 1;
constexpr int MODE_BUFFER = 2;

// ------------------------------------------------------------ CRS rows
// int32 x 4 with 4-byte alignment (gfx9 global stores take dword-aligned
// multi-dword accesses).
typedef int32_t i32x4_u __attribute__((ext_vector_type(4), aligned(4)));

// Result sink of one row.  MODE_COUNT only counts; MODE_BUFFER keeps the
// first `buffer` hits in the row's scratch slot; MODE_FILL writes the row of
// the CRS.  Hits leave in 16-byte groups (one store per 4 hits).
template <int MODE>
struct RowWriter
{
  int32_t *row;   // the row's first entry (unused when counting)
  uint32_t limit; // entries the row holds
  uint32_t cnt = 0;

  // Row q: FILL writes the final row at offsets[q], BUFFER a scratch row of
  // `buffer` entries, COUNT nothing.
  AXB_D void init(int32_t *out, const int64_t *offsets, int64_t q, int32_t buffer)
  {
    const int64_t base =
        MODE == MODE_FILL ? offsets[q] : (MODE == MODE_BUFFER ? q * (int64_t)buffer : 0);
    row = out + base;
    limit = MODE == MODE_FILL ? (uint32_t)(offsets[q + 1] - base) : (uint32_t)buffer;
  }
  // Hits are written in groups of four (one 16-byte store): the last three
  // wait in h0..h2, shifted along at every hit (the walk calls this in
  // nearly every step of a wavefront, some lane having a hit, so a hit
  // costs three moves and a compare).
  int32_t h0 = 0, h1 = 0, h2 = 0;
  AXB_D bool hit(int32_t v)
  {
    if (MODE != MODE_COUNT)
    {
      if ((cnt & 3u) == 3u && cnt < limit)
        *(i32x4_u *)(row + cnt - 3) = (i32x4_u){h0, h1, h2, v};
      h0 = h1, h1 = h2, h2 = v;
    }
    ++cnt;
    return true;
  }
  // Writes the last partial group of the hits (nothing for a row that
  // overflowed: its count sends it to be walked again).
  AXB_D void finish()
  {
    if (MODE == MODE_COUNT || cnt > limit)
      return;
    const uint32_t rem = cnt & 3u;
    int32_t *at = row + (cnt - rem);
    if (rem == 3)
      at[0] = h0, at[1] = h1, at[2] = h2;
    else if (rem == 2)
      at[0] = h1, at[1] = h2;
    else if (rem == 1)
      at[0] = h2;
  }
};

// Row writer of the spatial walks: hits gather in LDS, ROW_STAGE per lane
// (entry j of lane l at st[j * 256], st = the block's stage + l), and leave
// for the row as whole 64-byte chunks, written back to back (a row's 16-byte
// groups written at different steps left its lines to be written back
// partially, several times); the walk's loop carries only the count.
constexpr int ROW_STAGE = 16;
template <int MODE>
struct StagedRow
{
  int32_t *row;
  uint32_t limit;
  uint32_t cnt = 0;
  int32_t *st;

  AXB_D void init(int32_t *out, const int64_t *offsets, int64_t q, int32_t buffer,
                  int32_t *stage)
  {
    const int64_t base =
        MODE == MODE_FILL ? offsets[q] : (MODE == MODE_BUFFER ? q * (int64_t)buffer : 0);
    row = out + base;
    limit = MODE == MODE_FILL ? (uint32_t)(offsets[q + 1] - base) : (uint32_t)buffer;
    st = stage;
  }
  AXB_D bool hit(int32_t v)
  {
    if (MODE != MODE_COUNT)
    {
      const uint32_t j = cnt & (ROW_STAGE - 1);
      st[j * 256] = v;
      // the chunk [cnt - 15, cnt] is complete and inside the row
      if (j == ROW_STAGE - 1 && cnt < limit)
      {
        i32x4_u *dst = (i32x4_u *)(row + (cnt - (ROW_STAGE - 1)));
#pragma unroll
        for (int g = 0; g < ROW_STAGE / 4; ++g)
          dst[g] = (i32x4_u){st[(4 * g) * 256], st[(4 * g + 1) * 256], st[(4 * g + 2) * 256],
                             st[(4 * g + 3) * 256]};
      }
    }
    ++cnt;
    return true;
  }
  // Writes the last partial chunk (nothing for a row that overflowed: its
  // count sends it to be walked again).
  AXB_D void finish()
  {
    if (MODE == MODE_COUNT || cnt > limit)
      return;
    const uint32_t rem = cnt & (ROW_STAGE - 1);
    int32_t *at = row + (cnt - rem);
    for (uint32_t j = 0; j < rem; ++j)
      at[j] = st[j * 256];
  }
};

template <int MODE, template <int> class W>
AXB_D void row_epilogue(const TreeView &tv, W<MODE> &rw, int64_t *offsets, int64_t q,
                        uint32_t v_int, uint32_t v_leaf)
{
  rw.finish();
  if (MODE != MODE_FILL)
    offsets[q] = rw.cnt;
  else if (rw.cnt != rw.limit)
    raise_error(tv.err, KERR_LAYOUT);
  if (tv.visits && MODE != MODE_FILL)
  {
    atomicAdd(tv.visits + 0, 1ull);
    atomicAdd(tv.visits + 1, (unsigned long long)v_int);
    atomicAdd(tv.visits + 2, (unsigned long long)v_leaf);
    atomicAdd(tv.visits + 3, (unsigned long long)rw.cnt);
  }
}


constexpr int SLOT_PAD = 4; // zeroed slots after the array (the box walks read slot p + 1)
int64_t total_slots_for(int64_t n, int geometry, int node_format);
// Scene box {lo xyz, hi xyz} of n points or boxes (TreeConstruction.hpp:26-39
// semantics; the empty box {+FLT_MAX, -FLT_MAX} for n = 0) into device scene.
hipError_t scene_box(const float *values, int64_t n, int geometry, float *scene,
                     hipStream_t stream);
hipError_t build_tree(const float *values, int64_t n, int geometry, int node_format,
                      bool keep_topology, uint32_t *err, hipStream_t stream, BuildOut &out);
// Words of a build's status block: {error word, C[n - 1] (0 when n < 2),
// scene box (6 floats)}; build_status fills words 1..7 after build_tree.
constexpr int BUILD_STATUS_WORDS = 8;
hipError_t build_status(const BuildOut &b, int64_t n, uint32_t *status, hipStream_t stream);
hipError_t export_topology(const INode *inodes, int64_t nint, int32_t *left_child, int32_t *rope,
                           float *boxes, hipStream_t stream);
hipError_t launch_spatial(const TreeView &tv, int node_format, int geometry, int pred_kind,
                          const float *preds, int64_t nq, float radius, const uint32_t *order,
                          int64_t *offsets, int32_t *out, int32_t buffer, int mode,
                          hipStream_t s);
// Best-first ray queries over a box tree (rays nq x 6 {origin, unit
// direction}): row q holds the leaves ray q hits in (distance, index) order,
// at most `limit` of them.
hipError_t launch_ordered(const TreeView &tv, int node_format, const float *rays, int64_t nq,
                          int32_t limit, const uint32_t *order, int64_t *offsets, int32_t *out,
                          int32_t buffer, int mode, hipStream_t s);
// Intersects<Ray> rows holding the slots of the hit leaf records (box
// trees), and their conversion to rows of values in (distance, index)
// order: the rows of an ordered query without the best-first walk.
hipError_t launch_ray_slots(const TreeView &tv, int node_format, const float *rays, int64_t nq,
                            const uint32_t *order, int64_t *offsets, int32_t *out,
                            int32_t buffer, int mode, hipStream_t s);
// Whole rows of OrderedSpatial<Sphere | Box | Point> (pred_kind PRED_SPHERES,
// PRED_BOXES, PRED_POINTS): row q = rows[q * n, (q + 1) * n), every leaf's
// value in (distance, index) order.
hipError_t launch_ordered_rows(const TreeView &tv, int node_format, int geometry, int pred_kind,
                               const float *preds, int64_t nq, int32_t *rows, hipStream_t s);
hipError_t launch_ray_sort_rows(const TreeView &tv, const float *rays, int64_t nq,
                                const int64_t *offsets, int32_t *rows, hipStream_t s);
// Rows cut to their first k entries: counts[q] = min(k, row length) and
// counts[nq] = 0 (scanned into offsets by the caller), then the copy.
hipError_t launch_row_prefix_counts(const int64_t *offsets, int64_t nq, int32_t k,
                                    int64_t *counts, hipStream_t s);
hipError_t launch_row_prefix_copy(const int64_t *offsets, const int32_t *rows, int64_t nq,
                                  const int64_t *out_offsets, int32_t *out, hipStream_t s);
// Ray construction: rays[i] = {origins[i], directions[i] / |directions[i]|}
// with the norm in double; a zero direction raises KERR_LAYOUT in err.
hipError_t launch_make_rays(const float *origins, const float *directions, int64_t n,
                            float *rays, uint32_t *err, hipStream_t s);
hipError_t launch_knn(const TreeView &tv, int node_format, int geometry, const float *pts,
                      int64_t nq, int k, const uint32_t *order, int32_t *oi, float *od,
                      hipStream_t s);
// Buffers of a kNN call for k <= 64: candidate rows for `batch` queries
// (batch * knn_row_capacity(k, n) entries, batch * capacity + capacity
// below 2^32), the list of unresolved queries (nq entries) and its counter.
struct KnnScratch
{
  uint2 *cand;
  int64_t batch;
  uint32_t *ovf;
  uint32_t *ovf_n;
};
int knn_row_capacity(int k, int n);
hipError_t launch_knn_offsets(int64_t *offsets, int64_t nq, int64_t per, hipStream_t s);
// keys = the predicates' 30-bit Morton codes >> shift, and the histograms of
// their four 8-bit digits
hipError_t launch_pred_morton32(int pred_kind, const float *preds, int64_t nq,
                                const float *scene, int shift, uint32_t *keys, uint32_t *hist,
                                hipStream_t s);
hipError_t launch_compact(const int32_t *buf, int64_t nq, int32_t buffer,
                          const int64_t *offsets, int32_t *out, hipStream_t s);
hipError_t launch_overflow_list(const uint8_t *over_flags, const uint32_t *order, int64_t nq,
                                bool emit_index, uint32_t *ovf, uint32_t *ovf_n, hipStream_t s);
// Neighbour list of a point tree: lane t takes leaf position list[t] (t when
// list is nullptr); rows are the leaves' values.  rowof (leaf position ->
// row) is written when non-null.
hipError_t launch_nlist(const TreeView &tv, int node_format, bool full, float radius, int64_t nl,
                        const uint32_t *list, int64_t *offsets, int32_t *out, int32_t buffer,
                        uint32_t *rowof, int mode, hipStream_t s);
// Half -> full neighbour list: counts into cnt (zeroed, n + 1 entries), then
// after the scan into out_offsets, the fill (cursor: n scratch entries).
hipError_t launch_h2f_count(const int64_t *offsets, const int32_t *indices, int64_t n,
                            int64_t *cnt, uint32_t *err, hipStream_t s);
hipError_t launch_h2f_fill(const int64_t *offsets, const int32_t *indices, int64_t n,
                           const int64_t *out_offsets, int64_t *cursor, int32_t *out,
                           hipStream_t s);
// Brute-force index (every predicate against every value): spatial rows in
// value order through the RowWriter modes of the tree queries (lane t takes
// predicate list ? list[t] : t, nq of them); nearest rows of min(k, n)
// (distance, index) pairs, for k <= 64 by the register selection, above by
// a heap in the output row.
// Nearest<Ray> over a brute-force index of boxes: rows of min(k, hits)
// values by (distance, index), through the CRS row writers' modes.
hipError_t launch_bf_ray_nearest(const float *values, int64_t n, const float *rays, int64_t nq,
                                 int k, const uint32_t *list, int64_t *offsets, int32_t *out,
                                 int32_t buffer, int mode, uint32_t *err, hipStream_t s);
hipError_t launch_bf_spatial(const float *values, int64_t n, int geometry, int pred_kind,
                             const float *preds, int64_t nq, float radius, const uint32_t *list,
                             int64_t *offsets, int32_t *out, int32_t buffer, int mode,
                             uint32_t *err, hipStream_t s);
hipError_t launch_bf_nearest(const float *values, int64_t n, int geometry, const float *pts,
                             int64_t nq, int k, int32_t *oi, float *od, hipStream_t s);
// Rows are processed in `order` (nullptr: 0 .. n-1); E and grad do not
// depend on it except for the summation order of E.
hipError_t pair_energy(const float *xyz, int64_t n, const int64_t *offsets,
                       const int32_t *indices, float cutoff, const uint32_t *order,
                       double *energy, float *grad, hipStream_t s);
// The same over the library's static scratch of ENERGY_SCRATCH_BLOCKS
// partial sums (no allocation): n up to ENERGY_SCRATCH_BLOCKS * 256 rows,
// rows in the caller's order.
constexpr int64_t ENERGY_SCRATCH_BLOCKS = int64_t(1) << 19; // 4 MiB of partial sums
constexpr int MAX_DEVICES = 64;
hipError_t pair_energy_noalloc(const float *xyz, int64_t n, const int64_t *offsets,
                               const int32_t *indices, float cutoff, double *energy,
                               float *grad, hipStream_t s);
hipError_t pair_energy_fd(const float *xyz, const int64_t *offsets, const int32_t *indices,
                          float cutoff, const int32_t *atoms, int64_t m, double h, double *fd,
                          hipStream_t s);

// Bounded spin count for decoupled look-back waits.
constexpr uint32_t SPIN_LIMIT = 1u << 22;

// Atomic helpers at device (agent) scope.
AXB_D uint32_t ld_agent(const uint32_t *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
AXB_D void st_agent(uint32_t *p, uint32_t v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
AXB_D uint64_t ld_agent64(const uint64_t *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
AXB_D void st_agent64(uint64_t *p, uint64_t v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
} // namespace axb
