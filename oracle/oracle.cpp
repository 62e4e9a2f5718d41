// oracle.cpp -- CPU restatement of ArborX's Serial/OpenMP BVH hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library in
// arborx_amd/ is compared against.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it.  Nothing in the product path may
// call it.
//
// Built with g++ -O2 -ffp-contract=off -fopenmp (no FMA contraction, no
// fast-math), so every float expression below rounds exactly like the
// reference's Kokkos Serial/OpenMP build does.
//
// Every function cites the reference file:line it restates (paths relative to
// the ArborX source tree).  Deliberate differences from the reference:
//   * kNN ties at the k-th distance are broken by (distance, original index)
//     instead of traversal order (TreeTraversal.hpp:255,274,329 leave them
//     unspecified); pruning therefore uses `<=` where the reference uses `<`.
//   * predicate processing order (Morton32 sort of queries,
//     CrsGraphWrapperImpl.hpp:407-419) does not change any result row, so the
//     oracle processes queries in their original order.
//   * rays: a zero direction component is stored as +0 (orc_make_rays), and
//     the ordered (best-first) traversal stops when both children of the
//     node it stands on are missed and its heap is empty, where the
//     reference descends into the closer child anyway
//     (TreeTraversal.hpp:468-478) and so reports a leaf the ray misses.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc
{

// ---------------------------------------------------------------- geometry
// Box.hpp:34-44: an "empty" box is [+FLT_MAX, -FLT_MAX].
struct Box
{
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
  float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
};

// Kokkos::min / Kokkos::max as used by Expand.hpp:42-76.
static inline float kmin(float a, float b) { return (b < a) ? b : a; }
static inline float kmax(float a, float b) { return (a < b) ? b : a; }
// Kokkos::clamp (std::clamp semantics) as used by MortonCode.hpp:311,342.
template <class F>
static inline F kclamp(F v, F lo, F hi)
{
  return (v < lo) ? lo : ((hi < v) ? hi : v);
}

// Expand.hpp:42-57 (box += point) and :60-76 (box += box).
static inline void expand_point(Box &b, const float *p)
{
  for (int d = 0; d < 3; ++d)
  {
    b.lo[d] = kmin(b.lo[d], p[d]);
    b.hi[d] = kmax(b.hi[d], p[d]);
  }
}
static inline void expand_box(Box &b, const float *lo, const float *hi)
{
  for (int d = 0; d < 3; ++d)
  {
    b.lo[d] = kmin(b.lo[d], lo[d]);
    b.hi[d] = kmax(b.hi[d], hi[d]);
  }
}

// Distance.hpp:52-70: sequential sum from 0, then Kokkos::sqrt.
static inline float dist_pp(const float *a, const float *b)
{
  float s = 0;
  for (int d = 0; d < 3; ++d)
  {
    float t = b[d] - a[d];
    s += t * t;
  }
  return std::sqrt(s);
}
// ClosestPoint.hpp:48-66 then Distance.hpp:72-80 (distance point-box).
static inline float dist_pbox(const float *p, const float *lo, const float *hi)
{
  float c[3];
  for (int d = 0; d < 3; ++d)
  {
    if (p[d] < lo[d])
      c[d] = lo[d];
    else if (p[d] > hi[d])
      c[d] = hi[d];
    else
      c[d] = p[d];
  }
  return dist_pp(p, c);
}
// Intersects.hpp:51-65 (box-box).
static inline bool intersects_bb(const float *alo, const float *ahi,
                                 const float *blo, const float *bhi)
{
  for (int d = 0; d < 3; ++d)
    if (alo[d] > bhi[d] || ahi[d] < blo[d])
      return false;
  return true;
}

// Distance.hpp:159-195 (distance box-box): per axis the gap between the
// boxes when they are apart (a_min > b_max: a_min - b_max; b_min > a_max:
// b_min - a_max), squared and summed from 0 over the separated axes, then
// Kokkos::sqrt.
static inline float dist_bb(const float *alo, const float *ahi, const float *blo,
                            const float *bhi)
{
  float s = 0;
  for (int d = 0; d < 3; ++d)
  {
    if (alo[d] > bhi[d])
    {
      const float t = alo[d] - bhi[d];
      s += t * t;
    }
    else if (blo[d] > ahi[d])
    {
      const float t = blo[d] - ahi[d];
      s += t * t;
    }
  }
  return std::sqrt(s);
}
// Kokkos::max(a, b) = (a < b) ? b : a.
static inline float kokkos_max(float a, float b) { return (a < b) ? b : a; }
// Distance.hpp:196-208 (distance sphere-box): max(distance(centre, box) - r, 0).
static inline float dist_sbox(const float *c, float r, const float *lo, const float *hi)
{
  return kokkos_max(dist_pbox(c, lo, hi) - r, 0.f);
}
// Distance.hpp:82-94 (distance point-sphere): max(distance(p, centre) - r, 0).
static inline float dist_psphere(const float *p, const float *c, float r)
{
  return kokkos_max(dist_pp(p, c) - r, 0.f);
}

// ---------------------------------------------------------------- rays
// A ray is 6 floats {origin, unit direction} (ArborX_Ray.hpp:28-66).
// Slab intersection, ArborX_Ray.hpp:107-153.
static inline bool ray_interval(const float *ray, const float *lo,
                                const float *hi, float &tmin, float &tmax)
{
  const float inf = std::numeric_limits<float>::infinity();
  const float *orig = ray, *dir = ray + 3;
  tmin = -inf;
  tmax = inf;
  for (int d = 0; d < 3; ++d)
  {
    float tdmin, tdmax;
    if (dir[d] == 0)
    {
      const float min_orig = lo[d] - orig[d];
      if (min_orig == 0)
        continue;
      const float max_orig = hi[d] - orig[d];
      tdmin = std::signbit(dir[d] * max_orig) ? inf : -inf;
      tdmax = std::signbit(dir[d] * min_orig) ? inf : -inf;
    }
    else if (dir[d] > 0)
    {
      tdmin = (lo[d] - orig[d]) / dir[d];
      tdmax = (hi[d] - orig[d]) / dir[d];
    }
    else
    {
      tdmin = (hi[d] - orig[d]) / dir[d];
      tdmax = (lo[d] - orig[d]) / dir[d];
    }
    if (tmin < tdmin)
      tmin = tdmin;
    if (tmax > tdmax)
      tmax = tdmax;
  }
  return tmin <= tmax;
}
// ArborX_Ray.hpp:156-167: intersects(ray, box).
static inline bool ray_intersects(const float *ray, const float *lo,
                                  const float *hi)
{
  float tmin, tmax;
  return ray_interval(ray, lo, hi, tmin, tmax) && tmax >= 0;
}
// ArborX_Ray.hpp:429-444: distance(ray, box).
static inline float ray_distance(const float *ray, const float *lo,
                                 const float *hi)
{
  float tmin, tmax;
  const bool hit = ray_interval(ray, lo, hi, tmin, tmax) && tmax >= 0;
  return hit ? kmax(tmin, 0.f) : std::numeric_limits<float>::infinity();
}

// ---------------------------------------------------------------- Morton
// MortonCode.hpp:64-72 (expandBitsBy<2>, 32-bit).
static inline uint32_t expand2_u32(uint32_t x)
{
  x &= 0x000003ffu;
  x = (x ^ (x << 16)) & 0xff0000ffu;
  x = (x ^ (x << 8)) & 0x0300f00fu;
  x = (x ^ (x << 4)) & 0x030c30c3u;
  x = (x ^ (x << 2)) & 0x09249249u;
  return x;
}
// MortonCode.hpp:188-197 (expandBitsBy<2>, 64-bit).
static inline uint64_t expand2_u64(uint64_t x)
{
  x &= 0x1fffffull;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}
// MortonCode.hpp:291-316 for DIM=3: N = 2^10, computed in float.
static inline uint32_t morton32(const float *p)
{
  const unsigned N = 1u << 10;
  uint32_t r = 0;
  for (int d = 0; d < 3; ++d)
  {
    float x = kclamp(p[d] * (float)N, 0.f, (float)(N - 1));
    r += expand2_u32((uint32_t)x) << (3 - d - 1);
  }
  return r;
}
// MortonCode.hpp:318-347 for DIM=3: N = 2^21, NeedDouble false -> float.
static inline uint64_t morton64(const float *p)
{
  const unsigned long long N = 1ull << 21;
  uint64_t r = 0;
  for (int d = 0; d < 3; ++d)
  {
    float x = kclamp(p[d] * (float)N, 0.f, (float)(N - 1));
    r += expand2_u64((uint64_t)x) << (3 - d - 1);
  }
  return r;
}
// TranslateAndScale.hpp:23-37.
static inline void translate_and_scale(const float *in, float *out,
                                       const Box &ref)
{
  for (int d = 0; d < 3; ++d)
  {
    float a = ref.lo[d], b = ref.hi[d];
    out[d] = (a != b ? (in[d] - a) / (b - a) : 0);
  }
}
// Centroid.hpp:40-60: point -> itself, box -> (min+max)/2.
static inline void centroid(const float *v, int geometry, float *c)
{
  if (geometry == 0)
  {
    c[0] = v[0], c[1] = v[1], c[2] = v[2];
  }
  else
  {
    for (int d = 0; d < 3; ++d)
      c[d] = (v[d] + v[3 + d]) / 2;
  }
}

// ---------------------------------------------------------------- tree
// Node.hpp:24-51, HappyTreeFriends.hpp:23-86: leaves [0,n), internal nodes
// addressed as i+n, root = n, ROPE_SENTINEL = -1.
struct Tree
{
  int64_t n = 0;
  int geometry = 0; // 0 points (3 floats), 1 boxes (6 floats)
  Box bounds;
  std::vector<float> leaf_lo, leaf_hi;  // sorted leaves, 3 floats each
  std::vector<int32_t> leaf_rope;       // n
  std::vector<uint32_t> perm;           // n: original index of sorted leaf
  std::vector<uint64_t> keys;           // n: sorted Morton codes
  std::vector<int32_t> left_child;      // n-1
  std::vector<int32_t> rope;            // n-1
  std::vector<float> int_lo, int_hi;    // n-1 boxes
};

static inline int64_t delta(const std::vector<uint64_t> &k, int64_t n_int,
                            int64_t i)
{
  // TreeConstruction.hpp:131-170 (Apetrei delta with Karras tie-break).
  if (i < 0 || i >= n_int)
    return std::numeric_limits<int64_t>::max();
  uint64_t x = k[i] ^ k[i + 1];
  uint64_t r = x + (uint64_t)((!x) * (std::numeric_limits<int64_t>::min() +
                                      (int64_t)(i ^ (i + 1)))) -
               1;
  return (int64_t)r;
}

} // namespace orc

using namespace orc;

extern "C" {

// ------------------------------------------------------------- primitives
uint32_t orc_expand2_u32(uint32_t x) { return expand2_u32(x); }
uint64_t orc_expand2_u64(uint64_t x) { return expand2_u64(x); }
uint32_t orc_morton32(const float *p) { return morton32(p); }
uint64_t orc_morton64(const float *p) { return morton64(p); }

// TreeConstruction.hpp:26-39 + Reducer.hpp:22-67: scene bounding box.
void orc_scene_box(const float *values, int64_t n, int geometry, float *out6)
{
  Box b;
  int stride = geometry == 0 ? 3 : 6;
  for (int64_t i = 0; i < n; ++i)
  {
    const float *v = values + i * stride;
    if (geometry == 0)
      expand_point(b, v);
    else
      expand_box(b, v, v + 3);
  }
  for (int d = 0; d < 3; ++d)
    out6[d] = b.lo[d], out6[3 + d] = b.hi[d];
}

// SpaceFillingCurves.hpp:45-57,63-84: Morton64 of each value's centroid in
// the scene box.
void orc_morton64_values(const float *values, int64_t n, int geometry,
                         const float *scene6, uint64_t *out)
{
  Box s;
  for (int d = 0; d < 3; ++d)
    s.lo[d] = scene6[d], s.hi[d] = scene6[3 + d];
  int stride = geometry == 0 ? 3 : 6;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
  {
    float c[3], p[3];
    centroid(values + i * stride, geometry, c);
    translate_and_scale(c, p, s);
    out[i] = morton64(p);
  }
}
// SpaceFillingCurves.hpp:31-43: Morton32 of predicate centroids
// (geometry 0 point/sphere centre, 1 box).
void orc_morton32_values(const float *values, int64_t n, int stride,
                         int geometry, const float *scene6, uint32_t *out)
{
  Box s;
  for (int d = 0; d < 3; ++d)
    s.lo[d] = scene6[d], s.hi[d] = scene6[3 + d];
  for (int64_t i = 0; i < n; ++i)
  {
    float c[3], p[3];
    centroid(values + i * stride, geometry, c);
    translate_and_scale(c, p, s);
    out[i] = morton32(p);
  }
}

// SortUtils.hpp:26-43 + KokkosExtSort.hpp:79-149: stable sort of keys with
// the permutation of original indices.  Restated as a parallel LSD radix
// (8-bit digits): stable, so equal keys keep increasing original index,
// which is what Serial BinSort and the reference's Thrust path give.
void orc_sort_u64(uint64_t *keys, int64_t n, uint32_t *perm)
{
  std::vector<uint64_t> k2(n);
  std::vector<uint32_t> p2(n);
  for (int64_t i = 0; i < n; ++i)
    perm[i] = (uint32_t)i;
  uint64_t *ka = keys, *kb = k2.data();
  uint32_t *pa = perm, *pb = p2.data();
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  std::vector<int64_t> hist((size_t)nt * 256);
  int passes_done = 0;
  for (int pass = 0; pass < 8; ++pass)
  {
    int shift = pass * 8;
    std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(nt)
    {
      int t = 0;
#ifdef _OPENMP
      t = omp_get_thread_num();
#endif
      int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
      int64_t *h = &hist[(size_t)t * 256];
      for (int64_t i = lo; i < hi; ++i)
        h[(ka[i] >> shift) & 255]++;
    }
    // skip a pass in which every key has the same digit (identity)
    bool trivial = false;
    for (int dgt = 0; dgt < 256 && !trivial; ++dgt)
    {
      int64_t s = 0;
      for (int t = 0; t < nt; ++t)
        s += hist[(size_t)t * 256 + dgt];
      if (s == n)
        trivial = true;
    }
    if (trivial)
      continue;
    // exclusive offsets in (digit, thread) order => stable
    int64_t run = 0;
    for (int dgt = 0; dgt < 256; ++dgt)
      for (int t = 0; t < nt; ++t)
      {
        int64_t c = hist[(size_t)t * 256 + dgt];
        hist[(size_t)t * 256 + dgt] = run;
        run += c;
      }
#pragma omp parallel num_threads(nt)
    {
      int t = 0;
#ifdef _OPENMP
      t = omp_get_thread_num();
#endif
      int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
      int64_t *h = &hist[(size_t)t * 256];
      for (int64_t i = lo; i < hi; ++i)
      {
        int dgt = (ka[i] >> shift) & 255;
        int64_t o = h[dgt]++;
        kb[o] = ka[i];
        pb[o] = pa[i];
      }
    }
    std::swap(ka, kb);
    std::swap(pa, pb);
    ++passes_done;
  }
  if (ka != keys)
  {
    std::memcpy(keys, ka, n * sizeof(uint64_t));
    std::memcpy(perm, pa, n * sizeof(uint32_t));
  }
}
// std::stable_sort cross-check of the radix restatement.
void orc_stable_sort_u64(uint64_t *keys, int64_t n, uint32_t *perm)
{
  std::vector<uint32_t> idx(n);
  for (int64_t i = 0; i < n; ++i)
    idx[i] = (uint32_t)i;
  std::stable_sort(idx.begin(), idx.end(),
                   [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
  std::vector<uint64_t> k(n);
  for (int64_t i = 0; i < n; ++i)
    k[i] = keys[idx[i]], perm[i] = idx[i];
  std::memcpy(keys, k.data(), n * sizeof(uint64_t));
}

// ---------------------------------------------------------------- build
// LinearBVH.hpp:167-256: scene box -> Morton64 -> stable sort -> hierarchy.
static void *build_impl(const float *values, int64_t n, int geometry,
                        const uint64_t *given_sorted_codes)
{
  Tree *t = new Tree;
  t->n = n;
  t->geometry = geometry;
  if (n == 0)
    return t; // LinearBVH.hpp:203-206
  int stride = geometry == 0 ? 3 : 6;
  t->leaf_lo.resize(3 * n);
  t->leaf_hi.resize(3 * n);
  t->leaf_rope.assign(n, -1);
  t->perm.resize(n);
  t->keys.resize(n);
  if (n == 1)
  {
    // TreeConstruction.hpp:41-69: single leaf, bounds = its box.
    const float *v = values;
    Box b;
    if (geometry == 0)
      expand_point(b, v);
    else
      expand_box(b, v, v + 3);
    for (int d = 0; d < 3; ++d)
      t->leaf_lo[d] = b.lo[d], t->leaf_hi[d] = b.hi[d];
    t->bounds = b;
    t->perm[0] = 0;
    t->keys[0] = 0;
    return t;
  }
  if (given_sorted_codes)
  {
    // TreeConstruction tests drive generateHierarchy with explicit sorted
    // codes and an iota permutation (tstDetailsTreeConstruction.cpp:151-178).
    for (int64_t i = 0; i < n; ++i)
      t->keys[i] = given_sorted_codes[i], t->perm[i] = (uint32_t)i;
  }
  else
  {
    float scene[6];
    orc_scene_box(values, n, geometry, scene);
    orc_morton64_values(values, n, geometry, scene, t->keys.data());
    orc_sort_u64(t->keys.data(), n, t->perm.data());
  }

  const int64_t n_int = n - 1;
  t->left_child.assign(n_int, -1);
  t->rope.assign(n_int, -1);
  t->int_lo.assign(3 * n_int, 0.f);
  t->int_hi.assign(3 * n_int, 0.f);
  std::vector<std::atomic<int>> ranges(n_int);
  for (auto &r : ranges)
    r.store(-1, std::memory_order_relaxed);
  const auto &K = t->keys;

  // leaf boxes in sorted order (TreeConstruction.hpp:203-207)
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
  {
    const float *v = values + (int64_t)t->perm[i] * stride;
    Box b;
    if (geometry == 0)
      expand_point(b, v);
    else
      expand_box(b, v, v + 3);
    for (int d = 0; d < 3; ++d)
      t->leaf_lo[3 * i + d] = b.lo[d], t->leaf_hi[3 * i + d] = b.hi[d];
  }

  auto internal_index = [&](int64_t i) { return (int)(i + n_int + 1); };
  auto set_rope = [&](int range_right, int64_t delta_right) -> int {
    // TreeConstruction.hpp:172-195
    if (range_right != n_int)
      return delta_right < delta(K, n_int, range_right + 1)
                 ? range_right + 1
                 : internal_index(range_right + 1);
    return -1;
  };

  // TreeConstruction.hpp:197-311: bottom-up Apetrei climb, one leaf per
  // iteration; CAS flags make the first arrival stop.
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t leaf = 0; leaf < n; ++leaf)
  {
    int i = (int)leaf;
    Box bv;
    expand_box(bv, &t->leaf_lo[3 * leaf], &t->leaf_hi[3 * leaf]);
    int range_left = i, range_right = i;
    int64_t delta_left = delta(K, n_int, range_left - 1);
    int64_t delta_right = delta(K, n_int, range_right);
    t->leaf_rope[leaf] = set_rope(range_right, delta_right);
    const int root = internal_index(0);
    do
    {
      bool is_left_child = delta_right < delta_left;
      int left_child;
      if (is_left_child)
      {
        int apetrei_parent = range_right;
        int expected = -1;
        ranges[apetrei_parent].compare_exchange_strong(
            expected, range_left, std::memory_order_acq_rel);
        range_right = expected;
        if (range_right == -1)
          break;
        left_child = i;
        int right_child = apetrei_parent + 1;
        bool right_is_leaf = (right_child == range_right);
        delta_right = delta(K, n_int, range_right);
        if (right_is_leaf)
          expand_box(bv, &t->leaf_lo[3 * right_child],
                     &t->leaf_hi[3 * right_child]);
        else
          expand_box(bv, &t->int_lo[3 * right_child],
                     &t->int_hi[3 * right_child]);
      }
      else
      {
        int apetrei_parent = range_left - 1;
        int expected = -1;
        ranges[apetrei_parent].compare_exchange_strong(
            expected, range_right, std::memory_order_acq_rel);
        range_left = expected;
        if (range_left == -1)
          break;
        left_child = apetrei_parent;
        bool left_is_leaf = (left_child == range_left);
        delta_left = delta(K, n_int, range_left - 1);
        if (left_is_leaf)
          expand_box(bv, &t->leaf_lo[3 * left_child],
                     &t->leaf_hi[3 * left_child]);
        else
          expand_box(bv, &t->int_lo[3 * left_child],
                     &t->int_hi[3 * left_child]);
        if (!left_is_leaf)
          left_child = internal_index(left_child);
      }
      int karras_parent = delta_right < delta_left ? range_right : range_left;
      t->left_child[karras_parent] = left_child;
      t->rope[karras_parent] = set_rope(range_right, delta_right);
      for (int d = 0; d < 3; ++d)
        t->int_lo[3 * karras_parent + d] = bv.lo[d],
                                      t->int_hi[3 * karras_parent + d] =
                                          bv.hi[d];
      std::atomic_thread_fence(std::memory_order_release);
      i = internal_index(karras_parent);
    } while (i != root);
  }
  for (int d = 0; d < 3; ++d)
    t->bounds.lo[d] = t->int_lo[d], t->bounds.hi[d] = t->int_hi[d];
  return t;
}

void *orc_build(const float *values, int64_t n, int geometry)
{
  return build_impl(values, n, geometry, nullptr);
}
// Hierarchy over caller-given sorted codes, leaves = points at `values`
// (already in sorted order).
void *orc_build_from_codes(const uint64_t *sorted_codes, const float *values,
                           int64_t n)
{
  return build_impl(values, n, 0, sorted_codes);
}

void orc_free(void *h) { delete (Tree *)h; }
int64_t orc_size(void *h) { return ((Tree *)h)->n; }
void orc_bounds(void *h, float *out6)
{
  Tree *t = (Tree *)h;
  for (int d = 0; d < 3; ++d)
    out6[d] = t->bounds.lo[d], out6[3 + d] = t->bounds.hi[d];
}
// Export of the reference-form tree (Node.hpp layout) for topology parity.
void orc_export(void *h, uint64_t *keys, uint32_t *perm, int32_t *leaf_rope,
                int32_t *left_child, int32_t *rope, float *int_boxes6)
{
  Tree *t = (Tree *)h;
  int64_t n = t->n;
  if (n == 0)
    return;
  std::memcpy(keys, t->keys.data(), n * 8);
  std::memcpy(perm, t->perm.data(), n * 4);
  std::memcpy(leaf_rope, t->leaf_rope.data(), n * 4);
  if (n < 2)
    return;
  std::memcpy(left_child, t->left_child.data(), (n - 1) * 4);
  std::memcpy(rope, t->rope.data(), (n - 1) * 4);
  for (int64_t i = 0; i < n - 1; ++i)
    for (int d = 0; d < 3; ++d)
      int_boxes6[6 * i + d] = t->int_lo[3 * i + d],
                         int_boxes6[6 * i + 3 + d] = t->int_hi[3 * i + d];
}

// ------------------------------------------------- pre-order linearisation
// The GPU library stores the same hierarchy in depth-first pre-order in
// 16-byte slots: internal FP32 node = 2 slots {lo.xyz, skip}{hi.xyz, 0},
// internal Q16 node = 1 slot {6 x u16, skip}, point leaf = 1 slot
// {x,y,z,~orig}, box leaf = 2 slots {lo.xyz,~orig}{hi.xyz,0}.  The oracle
// derives it independently with an explicit DFS over left_child/rope
// (HappyTreeFriends.hpp:66-85: right child = rope of left child).
//
// Q16 encoding (no reference counterpart): a node's box is quantised on a
// 65535-step grid of a frame box F, one formula for both faces,
//   v(q) = fma(q, s, F.lo),  s = fma(F.hi - F.lo, C, 2^-149),
//   C = 0x1.00010ap-16
// (correctly rounded fused multiply-adds, as on the GPU; C > (1 + 2^-21) /
// 65535, so v(65535) >= F.hi; the smallest denormal keeps s > 0); qmin = largest q with v(q) <= box.lo, qmax =
// smallest q with v(q) >= box.hi (outward rounding => decoded box contains
// the exact one).  Frame = decoded box of the parent for a left child
// (reached by descent), the scene box for the root and every right child
// (reached by a rope).
static const float Q16_INV_UP = 0x1.00010ap-16f;

static inline float q16_v(float flo, float s, int q) { return std::fma((float)q, s, flo); }
static inline void q16_encode_axis(float flo, float fhi, float vlo, float vhi,
                                   uint16_t &qlo, uint16_t &qhi)
{
  float s = std::fma(fhi - flo, Q16_INV_UP, 0x1p-149f);
  if (!(s > 0.f))
  {
    qlo = 0, qhi = 65535;
    return;
  }
  // largest q with v(q) <= vlo (v(0) == flo <= vlo); v is monotone in q
  int a = 0, b = 65535;
  while (a < b)
  {
    int m = (a + b + 1) >> 1;
    if (q16_v(flo, s, m) <= vlo)
      a = m;
    else
      b = m - 1;
  }
  qlo = (uint16_t)a;
  // smallest q with v(q) >= vhi (v(65535) >= fhi >= vhi)
  a = 0, b = 65535;
  while (a < b)
  {
    int m = (a + b) >> 1;
    if (q16_v(flo, s, m) >= vhi)
      b = m;
    else
      a = m + 1;
  }
  qhi = (uint16_t)a;
}

void orc_q16_encode_box(const float *frame6, const float *box6, uint16_t *q6,
                        float *decoded6)
{
  for (int d = 0; d < 3; ++d)
  {
    uint16_t a, b;
    q16_encode_axis(frame6[d], frame6[3 + d], box6[d], box6[3 + d], a, b);
    q6[d] = a, q6[3 + d] = b;
    float s = std::fma(frame6[3 + d] - frame6[d], Q16_INV_UP, 0x1p-149f);
    decoded6[d] = (s > 0.f) ? q16_v(frame6[d], s, a) : frame6[d];
    decoded6[3 + d] = (s > 0.f) ? q16_v(frame6[d], s, b) : frame6[3 + d];
  }
}

// Writes the pre-order slot array (uint32 words, 4 per slot).  Returns the
// number of slots.  node_format 0 = FP32 internal nodes, 1 = Q16.
int64_t orc_linearize(void *h, int node_format, int flat_max, uint32_t *words)
{
  Tree *t = (Tree *)h;
  int64_t n = t->n;
  if (n == 0)
    return 0;
  const int s_leaf = t->geometry == 0 ? 1 : 2;
  const int s_int = node_format == 0 ? 2 : 1;
  auto put_f = [&](int64_t w, float f) { std::memcpy(&words[w], &f, 4); };
  auto write_leaf = [&](int64_t slot, int64_t leaf) {
    int64_t w = slot * 4;
    for (int d = 0; d < 3; ++d)
      put_f(w + d, t->leaf_lo[3 * leaf + d]);
    words[w + 3] = ~t->perm[leaf];
    if (s_leaf == 2)
    {
      for (int d = 0; d < 3; ++d)
        put_f(w + 4 + d, t->leaf_hi[3 * leaf + d]);
      words[w + 7] = 0;
    }
  };
  if (n == 1)
  {
    write_leaf(0, 0);
    return s_leaf;
  }
  float scene[6];
  for (int d = 0; d < 3; ++d)
    scene[d] = t->bounds.lo[d], scene[3 + d] = t->bounds.hi[d];
  // explicit DFS: (node, frame for Q16, subtree leaf count unknown -> compute)
  // subtree leaf counts via a post-order pass
  const int64_t N = n;
  std::vector<int64_t> len(n - 1, 0), first(n - 1, 0), nslots(n - 1, 0);
  auto right_of = [&](int32_t node) -> int32_t {
    int32_t lc = t->left_child[node - N];
    return lc < N ? t->leaf_rope[lc] : t->rope[lc - N];
  };
  {
    // iterative post-order
    std::vector<std::pair<int32_t, int>> st;
    st.push_back({(int32_t)N, 0});
    while (!st.empty())
    {
      auto &top = st.back();
      int32_t node = top.first;
      if (top.second == 0)
      {
        top.second = 1;
        int32_t lc = t->left_child[node - N];
        int32_t rc = right_of(node);
        if (rc >= N)
          st.push_back({rc, 0});
        if (lc >= N)
          st.push_back({lc, 0});
      }
      else
      {
        int32_t lc = t->left_child[node - N];
        int32_t rc = right_of(node);
        const int64_t k = node - N;
        len[k] = (lc < N ? 1 : len[lc - N]) + (rc < N ? 1 : len[rc - N]);
        first[k] = lc < N ? lc : first[lc - N];
        // a subtree of at most flat_max leaves is stored as its leaves only
        if (len[k] <= flat_max)
          nslots[k] = s_leaf * len[k];
        else
          nslots[k] = s_int + (lc < N ? s_leaf : nslots[lc - N]) + (rc < N ? s_leaf : nslots[rc - N]);
        st.pop_back();
      }
    }
  }
  struct Item
  {
    int32_t node;
    float frame[6];
  };
  std::vector<Item> st;
  Item root;
  root.node = (int32_t)N;
  std::memcpy(root.frame, scene, sizeof(scene));
  st.push_back(root);
  int64_t slot = 0;
  while (!st.empty())
  {
    Item it = st.back();
    st.pop_back();
    if (it.node < N)
    {
      write_leaf(slot, it.node);
      slot += s_leaf;
      continue;
    }
    int64_t k = it.node - N;
    if (len[k] <= flat_max)
    {
      for (int64_t i = first[k]; i < first[k] + len[k]; ++i)
      {
        write_leaf(slot, i);
        slot += s_leaf;
      }
      continue;
    }
    uint32_t skip = (uint32_t)nslots[k];
    int64_t w = slot * 4;
    float box[6];
    for (int d = 0; d < 3; ++d)
      box[d] = t->int_lo[3 * k + d], box[3 + d] = t->int_hi[3 * k + d];
    float dec[6];
    if (node_format == 0)
    {
      for (int d = 0; d < 3; ++d)
        put_f(w + d, box[d]), put_f(w + 4 + d, box[3 + d]);
      words[w + 3] = skip;
      words[w + 7] = 0;
      std::memcpy(dec, box, sizeof(dec));
    }
    else
    {
      uint16_t q[6];
      orc_q16_encode_box(it.frame, box, q, dec);
      words[w + 0] = (uint32_t)q[0] | ((uint32_t)q[1] << 16);
      words[w + 1] = (uint32_t)q[2] | ((uint32_t)q[3] << 16);
      words[w + 2] = (uint32_t)q[4] | ((uint32_t)q[5] << 16);
      words[w + 3] = skip;
    }
    slot += s_int;
    Item r, l;
    r.node = right_of(it.node);
    std::memcpy(r.frame, scene, sizeof(scene));
    l.node = t->left_child[k];
    std::memcpy(l.frame, dec, sizeof(dec));
    st.push_back(r);
    st.push_back(l);
  }
  return slot;
}

// ------------------------------------------------------- spatial queries
// Predicate kinds: 0 sphere {x,y,z,r} (Intersects<Sphere>), 1 box
// {lo.xyz, hi.xyz} (Intersects<Box>), 3 ray {origin, direction}
// (Intersects<Ray>, box leaves only).
static inline bool pred_leaf(int kind, const float *q, const float *llo,
                             const float *lhi, int geometry)
{
  if (kind == 3)
    return ray_intersects(q, llo, lhi);
  if (kind == 0)
  {
    // Intersects.hpp:105-114 (sphere-point) / :82-91 (sphere-box)
    float d = geometry == 0 ? dist_pp(q, llo) : dist_pbox(q, llo, lhi);
    return d <= q[3];
  }
  return intersects_bb(q, q + 3, llo, lhi); // Intersects.hpp:51-65
}
static inline bool pred_node(int kind, const float *q, const float *lo,
                             const float *hi)
{
  if (kind == 3)
    return ray_intersects(q, lo, hi);
  if (kind == 0)
    return dist_pbox(q, lo, hi) <= q[3]; // Intersects.hpp:82-91
  return intersects_bb(q, q + 3, lo, hi);
}

} // extern "C"

// TreeTraversal.hpp:80-119: stackless rope traversal of one predicate.
// Calls emit(original index) for each hit; counts visits.
template <class Emit>
static void traverse_spatial(const Tree &t, int kind, const float *q,
                             Emit emit, int64_t &v_int, int64_t &v_leaf)
{
  const int64_t N = t.n;
  if (N == 0)
    return;
  if (N == 1)
  {
    ++v_leaf;
    if (pred_leaf(kind, q, &t.leaf_lo[0], &t.leaf_hi[0], t.geometry))
      emit((int32_t)t.perm[0]);
    return;
  }
  int32_t node = (int32_t)N;
  do
  {
    if (node < N)
    {
      ++v_leaf;
      if (pred_leaf(kind, q, &t.leaf_lo[3 * node], &t.leaf_hi[3 * node],
                    t.geometry))
        emit((int32_t)t.perm[node]);
      node = t.leaf_rope[node];
    }
    else
    {
      ++v_int;
      int64_t k = node - N;
      node = pred_node(kind, q, &t.int_lo[3 * k], &t.int_hi[3 * k])
                 ? t.left_child[k]
                 : t.rope[k];
    }
  } while (node != -1);
}

extern "C" {

// CrsGraphWrapperImpl.hpp:145-323 (default policy, buffer_size 0): count
// pass, exclusive scan, fill pass; rows in original predicate order, each
// row in traversal order.  Returns nnz; indices must hold nnz entries when
// non-null (call twice: first with indices == nullptr to size).
int64_t orc_query_spatial(void *h, int kind, const float *preds, int64_t nq,
                          int64_t *offsets, int32_t *indices,
                          int64_t *visits_int, int64_t *visits_leaf)
{
  Tree *t = (Tree *)h;
  int stride = kind == 0 ? 4 : 6;
  std::vector<int64_t> counts(nq, 0);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < nq; ++i)
  {
    int64_t c = 0, vi = 0, vl = 0;
    traverse_spatial(*t, kind, preds + i * stride, [&](int32_t) { ++c; }, vi,
                     vl);
    counts[i] = c;
    if (visits_int)
      visits_int[i] = vi;
    if (visits_leaf)
      visits_leaf[i] = vl;
  }
  int64_t run = 0;
  for (int64_t i = 0; i < nq; ++i)
  {
    offsets[i] = run;
    run += counts[i];
  }
  offsets[nq] = run;
  if (indices)
  {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < nq; ++i)
    {
      int64_t o = offsets[i], vi = 0, vl = 0;
      traverse_spatial(*t, kind, preds + i * stride,
                       [&](int32_t v) { indices[o++] = v; }, vi, vl);
    }
  }
  return run;
}

// Brute-force O(N*M) spatial query with the same float formulas, rows in
// increasing original index (BruteForceImpl.hpp shape, used as a pin).
int64_t orc_bruteforce_spatial(const float *values, int64_t n, int geometry,
                               int kind, const float *preds, int64_t nq,
                               int64_t *offsets, int32_t *indices)
{
  int stride = kind == 0 ? 4 : 6;
  int vstride = geometry == 0 ? 3 : 6;
  std::vector<std::vector<int32_t>> rows(nq);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < nq; ++i)
  {
    const float *q = preds + i * stride;
    for (int64_t j = 0; j < n; ++j)
    {
      const float *v = values + j * vstride;
      const float *hi = geometry == 0 ? v : v + 3;
      if (pred_leaf(kind, q, v, hi, geometry))
        rows[i].push_back((int32_t)j);
    }
  }
  int64_t run = 0;
  for (int64_t i = 0; i < nq; ++i)
  {
    offsets[i] = run;
    if (indices)
      std::copy(rows[i].begin(), rows[i].end(), indices + run);
    run += rows[i].size();
  }
  offsets[nq] = run;
  return run;
}

// ------------------------------------------------------------------ kNN
struct Cand
{
  float d;
  int32_t idx; // original index
};
static inline bool cand_less(const Cand &a, const Cand &b)
{
  return a.d < b.d || (a.d == b.d && a.idx < b.idx);
}

// ------------------------------------------------------- ordered (rays)
// Ray construction, ArborX_Ray.hpp:42-51 with misc/ArborX_Vector.hpp:41-55
// and :77-87: norm = sqrt of a double sum r = 0; r += v_d * v_d; each
// component v_d / norm in double, rounded to float.  A zero component is
// stored as +0 (see the header).  Returns the number of zero directions
// (the reference asserts norm > 0); their rays are left as NaN.
int64_t orc_make_rays(const float *origins, const float *dirs, int64_t n,
                      float *out6)
{
  int64_t bad = 0;
  for (int64_t i = 0; i < n; ++i)
  {
    const float *v = dirs + 3 * i;
    double r = 0;
    for (int d = 0; d < 3; ++d)
      r += static_cast<double>(v[d]) * static_cast<double>(v[d]);
    const double mag = std::sqrt(r);
    bad += !(mag > 0);
    for (int d = 0; d < 3; ++d)
    {
      out6[6 * i + d] = origins[3 * i + d];
      float w = mag > 0 ? (float)(v[d] / mag) : std::nanf("");
      out6[6 * i + 3 + d] = w == 0.f ? 0.f : w;
    }
  }
  return bad;
}

// OrderedSpatial<G>::distance(other) = distance(G, other)
// (Predicates.hpp:104-126) for the predicate kinds of the C ABI: 4 rays
// {origin, direction} (trees of boxes), 5 spheres {centre, r}, 6 boxes
// {lo, hi}, 7 points.  `other` is a leaf's value (a point when `point`,
// lo = the point) or a box; pairs the reference defines one way round only
// go through its reverse dispatch (ReverseDispatch.hpp:32-46):
// distance(sphere, point) = distance(point, sphere), distance(box, point)
// = distance(point, box).
static inline int ordered_stride(int kind)
{
  return kind == 5 ? 4 : kind == 7 ? 3 : 6;
}
static inline float ordered_dist(int kind, const float *q, bool point, const float *lo,
                                 const float *hi)
{
  switch (kind)
  {
  case 4:
    return ray_distance(q, lo, hi);
  case 5:
    return point ? dist_psphere(lo, q, q[3]) : dist_sbox(q, q[3], lo, hi);
  case 6:
    return point ? dist_pbox(lo, q, q + 3) : dist_bb(q, q + 3, lo, hi);
  default:
    return point ? dist_pp(q, lo) : dist_pbox(q, lo, hi);
  }
}
float orc_ordered_distance(int kind, const float *pred, int geometry, const float *value)
{
  return ordered_dist(kind, pred, geometry == 0, value, geometry == 0 ? value : value + 3);
}

// Ordered rows by brute force: every value at a finite distance (the
// traversal below never pushes +inf; rays: the boxes the ray hits, other
// geometries: every value), sorted by (distance, index), at most `limit`
// per row (limit < 0: all).  For rays this is also the row of a
// Nearest<Ray> query with k = limit: the kNN traversal only keeps leaves at
// a distance below its radius, which starts at +inf
// (TreeTraversal.hpp:255-275), so boxes the ray misses never qualify.
int64_t orc_bruteforce_ordered(const float *values, int64_t n, int geometry, int kind,
                               const float *preds, int64_t nq, int64_t limit,
                               int64_t *offsets, int32_t *indices)
{
  const int vs = geometry == 0 ? 3 : 6, ps = ordered_stride(kind);
  std::vector<std::vector<int32_t>> rows(nq);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < nq; ++i)
  {
    std::vector<Cand> hits;
    for (int64_t j = 0; j < n; ++j)
    {
      const float *v = values + vs * j;
      const float d = ordered_dist(kind, preds + ps * i, geometry == 0, v,
                                   geometry == 0 ? v : v + 3);
      if (d < std::numeric_limits<float>::infinity())
        hits.push_back({d, (int32_t)j});
    }
    std::sort(hits.begin(), hits.end(), cand_less);
    const size_t m = limit < 0 ? hits.size() : std::min<size_t>(hits.size(), (size_t)limit);
    for (size_t j = 0; j < m; ++j)
      rows[i].push_back(hits[j].idx);
  }
  int64_t run = 0;
  for (int64_t i = 0; i < nq; ++i)
  {
    offsets[i] = run;
    if (indices)
      std::copy(rows[i].begin(), rows[i].end(), indices + run);
    run += rows[i].size();
  }
  offsets[nq] = run;
  return run;
}
int64_t orc_bruteforce_rays_ordered(const float *boxes, int64_t n, const float *rays,
                                    int64_t nq, int64_t limit, int64_t *offsets,
                                    int32_t *indices)
{
  return orc_bruteforce_ordered(boxes, n, 1, 4, rays, nq, limit, offsets, indices);
}

// TreeTraversal.hpp:383-489: the ordered spatial traversal over the
// reference-form tree.  OneLeafTree: the leaf if its distance is finite.
// FullTree: from the root, a leaf is reported and the heap's top taken; at
// an internal node the distances of both children are computed, the closer
// (distance_left < distance_right ? left : right) is taken unless the heap's
// top is closer (then the top is taken and the closer child pushed), and
// the further child is pushed, children at +inf never.  The heap
// (PriorityQueue with CompareDistance, 64 entries) is keyed by distance
// alone, so equal distances come out in heap order; the rows are compared
// after sorting equal distances by index.  The stop when both children are
// missed and the heap is empty is the difference listed in the header.
// Returns nnz, or -1 if a heap would exceed 64 entries.
int64_t orc_query_ordered_kind(void *h, int kind, const float *preds, int64_t nq,
                               int64_t *offsets, int32_t *indices, float *distances)
{
  const Tree &t = *(const Tree *)h;
  const int ps = ordered_stride(kind);
  const bool point_leaves = t.geometry == 0;
  const int64_t N = t.n;
  const float inf = std::numeric_limits<float>::infinity();
  std::vector<std::vector<std::pair<int32_t, float>>> rows(nq);
  std::atomic<bool> overflow{false};
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < nq; ++i)
  {
    const float *q = preds + ps * i;
    auto &row = rows[i];
    if (N == 0)
      continue;
    auto dist = [&](int64_t node) {
      return node < N ? ordered_dist(kind, q, point_leaves, &t.leaf_lo[3 * node],
                                     &t.leaf_hi[3 * node])
                      : ordered_dist(kind, q, false, &t.int_lo[3 * (node - N)],
                                     &t.int_hi[3 * (node - N)]);
    };
    if (N == 1)
    {
      const float d = dist(0);
      if (d != inf)
        row.push_back({(int32_t)t.perm[0], d});
      continue;
    }
    using PID = std::pair<int64_t, float>;
    auto further_first = [](const PID &a, const PID &b) { return a.second > b.second; };
    std::vector<PID> heap;
    int64_t node = N; // root
    while (true)
    {
      if (node < N)
      {
        row.push_back({(int32_t)t.perm[node], dist(node)});
        if (heap.empty())
          break;
        std::pop_heap(heap.begin(), heap.end(), further_first);
        node = heap.back().first;
        heap.pop_back();
        continue;
      }
      const int64_t l = t.left_child[node - N];
      const int64_t r = l < N ? t.leaf_rope[l] : t.rope[l - N];
      const PID lp{l, dist(l)}, rp{r, dist(r)};
      const PID &closer = lp.second < rp.second ? lp : rp;
      const PID &further = lp.second < rp.second ? rp : lp;
      if (!heap.empty() && heap.front().second < closer.second)
      {
        std::pop_heap(heap.begin(), heap.end(), further_first);
        node = heap.back().first;
        heap.pop_back();
        if (closer.second < inf)
        {
          heap.push_back(closer);
          std::push_heap(heap.begin(), heap.end(), further_first);
        }
      }
      else if (closer.second < inf)
        node = closer.first;
      else
        break; // both children missed, heap empty
      if (further.second < inf)
      {
        heap.push_back(further);
        std::push_heap(heap.begin(), heap.end(), further_first);
      }
      if (heap.size() > 64)
      {
        overflow = true;
        break;
      }
    }
  }
  if (overflow)
    return -1;
  int64_t run = 0;
  for (int64_t i = 0; i < nq; ++i)
  {
    offsets[i] = run;
    for (size_t j = 0; j < rows[i].size(); ++j)
    {
      if (indices)
        indices[run + j] = rows[i][j].first;
      if (distances)
        distances[run + j] = rows[i][j].second;
    }
    run += rows[i].size();
  }
  offsets[nq] = run;
  return run;
}

int64_t orc_query_ordered(void *h, const float *rays, int64_t nq, int64_t *offsets,
                          int32_t *indices, float *distances)
{
  return orc_query_ordered_kind(h, 4, rays, nq, offsets, indices, distances);
}

// TreeTraversal.hpp:168-335: stack-based DFS, closest child first, bounded
// max-heap (Heap.hpp / PriorityQueue.hpp: push, popPush, sortHeap) whose
// order is (distance, original index).  Output row sorted ascending.
static int knn_one(const Tree &t, const float *q, int k, Cand *out,
                   int64_t &v_int, int64_t &v_leaf)
{
  const int64_t N = t.n;
  if (N == 0 || k < 1)
    return 0;
  auto leaf_dist = [&](int64_t leaf) {
    return t.geometry == 0
               ? dist_pp(q, &t.leaf_lo[3 * leaf])
               : dist_pbox(q, &t.leaf_lo[3 * leaf], &t.leaf_hi[3 * leaf]);
  };
  if (N == 1)
  {
    ++v_leaf;
    out[0] = {leaf_dist(0), (int32_t)t.perm[0]};
    return 1;
  }
  std::vector<Cand> heap;
  heap.reserve(k);
  auto dist = [&](int32_t j) -> float {
    if (j < N)
    {
      ++v_leaf;
      return leaf_dist(j);
    }
    ++v_int;
    int64_t kk = j - N;
    return dist_pbox(q, &t.int_lo[3 * kk], &t.int_hi[3 * kk]);
  };
  float radius = std::numeric_limits<float>::infinity();
  auto insert = [&](int32_t leaf, float d) {
    Cand c{d, (int32_t)t.perm[leaf]};
    if ((int)heap.size() < k)
    {
      heap.push_back(c);
      std::push_heap(heap.begin(), heap.end(), cand_less);
    }
    else if (cand_less(c, heap.front()))
    {
      std::pop_heap(heap.begin(), heap.end(), cand_less);
      heap.back() = c;
      std::push_heap(heap.begin(), heap.end(), cand_less);
    }
    if ((int)heap.size() == k)
      radius = heap.front().d;
  };
  int32_t stack[128];
  float stack_d[128];
  int sp = 0;
  stack[sp] = -1, stack_d[sp] = 0.f, ++sp;
  int32_t node = (int32_t)N;
  float dnode = 0.f;
  do
  {
    bool tl = false, tr = false;
    int32_t lc = -1, rc = -1;
    float dl = 0.f, dr = 0.f;
    if (dnode <= radius)
    {
      lc = t.left_child[node - N];
      rc = lc < N ? t.leaf_rope[lc] : t.rope[lc - N];
      dl = dist(lc);
      dr = dist(rc);
      if (dl <= radius)
      {
        if (lc < N)
          insert(lc, dl);
        else
          tl = true;
      }
      if (dr <= radius)
      {
        if (rc < N)
          insert(rc, dr);
        else
          tr = true;
      }
    }
    if (!tl && !tr)
    {
      --sp;
      node = stack[sp];
      dnode = stack_d[sp];
    }
    else
    {
      node = (tl && (dl <= dr || !tr)) ? lc : rc;
      dnode = (node == lc ? dl : dr);
      if (tl && tr)
      {
        stack[sp] = (node == lc ? rc : lc);
        stack_d[sp] = (node == lc ? dr : dl);
        ++sp;
      }
    }
  } while (node != -1);
  std::sort(heap.begin(), heap.end(), cand_less);
  for (size_t i = 0; i < heap.size(); ++i)
    out[i] = heap[i];
  return (int)heap.size();
}

// CrsGraphWrapperImpl.hpp:353-374 + :296-318: rows of min(k, n) results.
int64_t orc_query_knn(void *h, const float *pts, int64_t nq, int k,
                      int64_t *offsets, int32_t *indices, float *dists,
                      int64_t *visits_int, int64_t *visits_leaf)
{
  Tree *t = (Tree *)h;
  int64_t per = (k < 1 || t->n == 0) ? 0 : std::min<int64_t>(k, t->n);
  for (int64_t i = 0; i <= nq; ++i)
    offsets[i] = i * per;
#pragma omp parallel
  {
    std::vector<Cand> buf(per > 0 ? per : 1);
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < nq; ++i)
    {
      int64_t vi = 0, vl = 0;
      int m = knn_one(*t, pts + 3 * i, k, buf.data(), vi, vl);
      for (int j = 0; j < m; ++j)
      {
        indices[i * per + j] = buf[j].idx;
        if (dists)
          dists[i * per + j] = buf[j].d;
      }
      if (visits_int)
        visits_int[i] = vi;
      if (visits_leaf)
        visits_leaf[i] = vl;
    }
  }
  return nq * per;
}

// Brute-force kNN with the (distance, index) order.
void orc_bruteforce_knn(const float *values, int64_t n, int geometry,
                        const float *pts, int64_t nq, int k, int32_t *indices,
                        float *dists)
{
  int64_t per = (k < 1 || n == 0) ? 0 : std::min<int64_t>(k, n);
  int vstride = geometry == 0 ? 3 : 6;
#pragma omp parallel
  {
    std::vector<Cand> all(n);
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < nq; ++i)
    {
      const float *q = pts + 3 * i;
      for (int64_t j = 0; j < n; ++j)
      {
        const float *v = values + j * vstride;
        float d = geometry == 0 ? dist_pp(q, v) : dist_pbox(q, v, v + 3);
        all[j] = {d, (int32_t)j};
      }
      std::partial_sort(all.begin(), all.begin() + per, all.end(), cand_less);
      for (int64_t j = 0; j < per; ++j)
      {
        indices[i * per + j] = all[j].idx;
        if (dists)
          dists[i * per + j] = all[j].d;
      }
    }
  }
}

// --------------------------------------------------------------- energy
// Pair energy over a fixed neighbour CSR (no reference counterpart; the
// closest precedent is examples/molecular_dynamics/
// example_molecular_dynamics.cpp:128-194).  Shifted Lennard-Jones
// phi(r) = 4(r^-12 - r^-6) - phi(rc); E = 1/2 sum_i sum_{j in row i, j != i}
// phi(r_ij); grad_i = sum_{j in row i, j != i} phi'(r)/r (x_i - x_j).
// Everything in FP64 here.
double orc_pair_energy_f64(const float *xyz, int64_t n, const int64_t *offsets,
                           const int32_t *indices, double rc, double *grad)
{
  double irc2 = 1.0 / (rc * rc);
  double irc6 = irc2 * irc2 * irc2;
  double phic = 4.0 * irc6 * (irc6 - 1.0);
  double e = 0.0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : e)
  for (int64_t i = 0; i < n; ++i)
  {
    double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
    double gx = 0, gy = 0, gz = 0, ei = 0;
    for (int64_t p = offsets[i]; p < offsets[i + 1]; ++p)
    {
      int32_t j = indices[p];
      if (j == i)
        continue;
      double dx = xi - xyz[3 * (int64_t)j], dy = yi - xyz[3 * (int64_t)j + 1],
             dz = zi - xyz[3 * (int64_t)j + 2];
      double r2 = dx * dx + dy * dy + dz * dz;
      double ir2 = 1.0 / r2;
      double ir6 = ir2 * ir2 * ir2;
      ei += 4.0 * ir6 * (ir6 - 1.0) - phic;
      double f = 24.0 * ir2 * ir6 * (1.0 - 2.0 * ir6); // phi'(r)/r
      gx += f * dx, gy += f * dy, gz += f * dz;
    }
    e += 0.5 * ei;
    if (grad)
      grad[3 * i] = gx, grad[3 * i + 1] = gy, grad[3 * i + 2] = gz;
  }
  return e;
}

} // extern "C"

// ------------------------------------------------------- neighbour lists
// HalfTraversal.hpp:45-72: for every leaf i (in leaf order) with the sphere
// {point_i, r}, walk from rope(i) to the end; emit(value_i, value_node) for
// each leaf hit.  Pairs are therefore (earlier leaf, later leaf).
template <class Emit>
static void half_traversal(const Tree &t, float r, Emit emit)
{
  const int64_t N = t.n;
  if (N <= 1) // HalfTraversal.hpp:33-44: nothing for 0 or 1 values
    return;
  for (int64_t i = 0; i < N; ++i)
  {
    float q[4] = {t.leaf_lo[3 * i], t.leaf_lo[3 * i + 1], t.leaf_lo[3 * i + 2], r};
    int32_t node = t.leaf_rope[i];
    while (node != -1)
    {
      if (node < N)
      {
        if (pred_leaf(0, q, &t.leaf_lo[3 * node], &t.leaf_hi[3 * node], 0))
          emit((int32_t)t.perm[i], (int32_t)t.perm[node]);
        node = t.leaf_rope[node];
      }
      else
      {
        int64_t k = node - N;
        node = pred_node(0, q, &t.int_lo[3 * k], &t.int_hi[3 * k]) ? t.left_child[k]
                                                                   : t.rope[k];
      }
    }
  }
}

extern "C" {

// NeighborList.hpp:49-103 (findHalfNeighborList) and :105-188
// (findFullNeighborList), serial.  offsets has n + 1 entries; indices is
// written when non-null (call twice: first with nullptr to size).  Row j of
// the half list holds the values i whose leaf comes before j's leaf; the
// full list holds both directions.  Returns nnz.
int64_t orc_neighbor_list(const float *points, int64_t n, float r, int full,
                          int64_t *offsets, int32_t *indices)
{
  std::fill(offsets, offsets + n + 1, 0);
  if (n == 0)
    return 0;
  Tree *t = (Tree *)build_impl(points, n, 0, nullptr);
  half_traversal(*t, r, [&](int32_t v1, int32_t v2) {
    if (full)
      ++offsets[v1];
    ++offsets[v2];
  });
  int64_t run = 0;
  for (int64_t i = 0; i <= n; ++i)
  {
    int64_t c = offsets[i];
    offsets[i] = run;
    run += c;
  }
  if (indices)
  {
    std::vector<int64_t> counts(offsets, offsets + n);
    half_traversal(*t, r, [&](int32_t v1, int32_t v2) { indices[counts[v2]++] = v1; });
    if (full)
    {
      // NeighborList.hpp:160-183: mirror each (row i, entry k) into row k
      std::vector<int64_t> last(counts);
      for (int64_t i = 0; i < n; ++i)
        for (int64_t p = offsets[i]; p < last[i]; ++p)
        {
          int32_t k = indices[p];
          indices[counts[k]++] = (int32_t)i;
        }
    }
  }
  delete t;
  return run;
}

// ExpandHalfToFull.hpp:23-68, serial: rows of the full list are the row's
// own entries plus the mirrored ones.  out_offsets n + 1; out_indices
// 2 * nnz(half) entries when non-null.
int64_t orc_expand_half_to_full(const int64_t *offsets, const int32_t *indices, int64_t n,
                                int64_t *out_offsets, int32_t *out_indices)
{
  std::fill(out_offsets, out_offsets + n + 1, 0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t p = offsets[i]; p < offsets[i + 1]; ++p)
    {
      ++out_offsets[i];
      ++out_offsets[indices[p]];
    }
  int64_t run = 0;
  for (int64_t i = 0; i <= n; ++i)
  {
    int64_t c = out_offsets[i];
    out_offsets[i] = run;
    run += c;
  }
  if (out_indices)
  {
    std::vector<int64_t> counts(out_offsets, out_offsets + n);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t p = offsets[i]; p < offsets[i + 1]; ++p)
      {
        int32_t k = indices[p];
        out_indices[counts[i]++] = k;
        out_indices[counts[k]++] = (int32_t)i;
      }
  }
  return run;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void orc_set_num_threads(int n)
{
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

} // extern "C"
