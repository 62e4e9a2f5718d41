// test_walk_host.cpp -- the library's traversal code (rope_walk and the
// best-first ordered_walk of include/arborx_mi355x_device.hpp, the
// functions the GPU kernels run) run on
// the CPU over node arrays written by the oracle (oracle/oracle.cpp), and
// compared with the oracle's reference-form rope traversal
// (TreeTraversal.hpp:97-119): identical rows in identical order, for FP32
// and Q16 nodes, flattened small subtrees, leaf groups, a skipped leaf and
// a shortened walk (the neighbour-list forms).  Host code only (hipcc host
// pass); no GPU is used.
#include "arborx_mi355x_device.hpp"

#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

extern "C" {
void *orc_build(const float *values, int64_t n, int geometry);
void orc_free(void *h);
int64_t orc_linearize(void *h, int node_format, int flat_max, uint32_t *words);
int64_t orc_query_spatial(void *h, int kind, const float *preds, int64_t nq, int64_t *offsets,
                          int32_t *indices, int64_t *visits_int, int64_t *visits_leaf);
void orc_bounds(void *h, float *out6);
int64_t orc_make_rays(const float *origins, const float *dirs, int64_t n, float *out6);
int64_t orc_bruteforce_rays_ordered(const float *boxes, int64_t n, const float *rays, int64_t nq,
                                    int64_t limit, int64_t *offsets, int32_t *indices);
int64_t orc_bruteforce_spatial(const float *values, int64_t n, int geometry, int kind,
                               const float *preds, int64_t nq, int64_t *offsets,
                               int32_t *indices);
}

static int g_fail = 0, g_pass = 0;
#define CHECK(cond, what)                                                                   \
  do                                                                                        \
  {                                                                                         \
    if (cond)                                                                               \
      ++g_pass;                                                                             \
    else                                                                                    \
    {                                                                                       \
      ++g_fail;                                                                             \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, what);                             \
    }                                                                                       \
  } while (0)

struct HostTree
{
  const uint4 *slots;
  axb::Box6 scene;
  float ss[3];
  uint32_t *err;
};

struct RowSink
{
  std::vector<int32_t> *row;
  bool hit(int32_t v)
  {
    row->push_back(v);
    return true;
  }
};

// One case: n points (geometry 0) or boxes (1), spheres around the first
// nq values of radius r.
static void run_case(int geometry, int64_t n, unsigned seed, float r, int node_format)
{
  std::mt19937 gen(seed);
  std::uniform_real_distribution<float> U(0.f, 1.f);
  const int vs = geometry == 0 ? 3 : 6;
  std::vector<float> v(n * vs);
  for (int64_t i = 0; i < n; ++i)
  {
    float x = U(gen), y = U(gen), z = U(gen);
    if (i % 17 == 5 && i > 0) // exact duplicates
      x = v[vs * (i - 1)], y = v[vs * (i - 1) + 1], z = v[vs * (i - 1) + 2];
    v[vs * i] = x, v[vs * i + 1] = y, v[vs * i + 2] = z;
    if (geometry == 1)
    {
      float e = 0.004f * U(gen);
      v[vs * i + 3] = x + e, v[vs * i + 4] = y + e, v[vs * i + 5] = z + e;
    }
  }
  void *t = orc_build(v.data(), n, geometry);
  std::vector<uint32_t> words((size_t)(n * 2 + n * 2 + 16) * 4, 0u);
  const int64_t total = orc_linearize(t, node_format, AXB_FLAT_SUBTREE_MAX, words.data());
  // the library keeps zeroed slots after the array for the leaf-group reads
  std::vector<uint4> slots((size_t)total + 8, make_uint4(0u, 0u, 0u, 0u));
  std::memcpy(slots.data(), words.data(), (size_t)total * sizeof(uint4));
  HostTree tv;
  tv.slots = slots.data();
  float b[6];
  orc_bounds(t, b);
  for (int d = 0; d < 3; ++d)
  {
    tv.scene.lo[d] = b[d], tv.scene.hi[d] = b[3 + d];
    tv.ss[d] = axb::q16_scale(b[d], b[3 + d]);
  }
  uint32_t err = 0;
  tv.err = &err;

  const int64_t nq = n < 400 ? n : 400;
  std::vector<float> sph(nq * 4);
  for (int64_t q = 0; q < nq; ++q)
  {
    for (int d = 0; d < 3; ++d)
      sph[4 * q + d] = v[vs * q + d];
    sph[4 * q + 3] = r;
  }
  std::vector<int64_t> off(nq + 1);
  const int64_t nnz = orc_query_spatial(t, 0, sph.data(), nq, off.data(), nullptr, nullptr,
                                        nullptr);
  std::vector<int32_t> ref(nnz > 0 ? nnz : 1);
  orc_query_spatial(t, 0, sph.data(), nq, off.data(), ref.data(), nullptr, nullptr);

  bool ok = true, ok_skip = true;
  for (int64_t q = 0; q < nq && ok; ++q)
  {
    const axb::SpatialPred pr =
        axb::sphere_pred(sph[4 * q], sph[4 * q + 1], sph[4 * q + 2], sph[4 * q + 3]);
    std::vector<int32_t> row;
    RowSink sink{&row};
    uint32_t vi = 0, vl = 0;
    if (geometry == 0)
    {
      if (node_format == 0)
        axb::rope_walk<axb::NODES_FP32, axb::GEOM_POINTS, axb::PRED_SPHERES>(
            tv, pr, 0u, (uint32_t)total, ~0u, sink, vi, vl);
      else
        axb::rope_walk<axb::NODES_Q16, axb::GEOM_POINTS, axb::PRED_SPHERES>(
            tv, pr, 0u, (uint32_t)total, ~0u, sink, vi, vl);
    }
    else
    {
      if (node_format == 0)
        axb::rope_walk<axb::NODES_FP32, axb::GEOM_BOXES, axb::PRED_SPHERES>(
            tv, pr, 0u, (uint32_t)total, ~0u, sink, vi, vl);
      else
        axb::rope_walk<axb::NODES_Q16, axb::GEOM_BOXES, axb::PRED_SPHERES>(
            tv, pr, 0u, (uint32_t)total, ~0u, sink, vi, vl);
    }
    ok = row.size() == (size_t)(off[q + 1] - off[q]) &&
         std::equal(row.begin(), row.end(), ref.begin() + off[q]);
    if (geometry == 0 && ok)
    {
      // the neighbour-list forms: skip this query's own leaf record, or
      // stop at it (the half list)
      uint32_t self = ~0u;
      for (int64_t p = 0; p < total; ++p)
        if ((int32_t)slots[p].w == ~(int32_t)q)
          self = (uint32_t)p;
      std::vector<int32_t> full, half;
      RowSink sf{&full}, sh{&half};
      if (node_format == 0)
      {
        axb::rope_walk<axb::NODES_FP32, axb::GEOM_POINTS, axb::PRED_SPHERES>(
            tv, pr, 0u, (uint32_t)total, self, sf, vi, vl);
        axb::rope_walk<axb::NODES_FP32, axb::GEOM_POINTS, axb::PRED_SPHERES>(
            tv, pr, 0u, self, self, sh, vi, vl);
      }
      else
      {
        axb::rope_walk<axb::NODES_Q16, axb::GEOM_POINTS, axb::PRED_SPHERES>(
            tv, pr, 0u, (uint32_t)total, self, sf, vi, vl);
        axb::rope_walk<axb::NODES_Q16, axb::GEOM_POINTS, axb::PRED_SPHERES>(
            tv, pr, 0u, self, self, sh, vi, vl);
      }
      std::vector<int32_t> want_full, want_half;
      bool before = true;
      for (int64_t j = off[q]; j < off[q + 1]; ++j)
      {
        if (ref[j] == (int32_t)q)
        {
          before = false;
          continue;
        }
        want_full.push_back(ref[j]);
        if (before)
          want_half.push_back(ref[j]);
      }
      ok_skip = ok_skip && full == want_full && half == want_half;
    }
  }
  char what[160];
  std::snprintf(what, sizeof what, "rows: geometry %d n %lld format %d", geometry, (long long)n,
                node_format);
  CHECK(ok && err == 0, what);

  // box predicates: boxes around the same centres, of three sizes; every
  // fifth has its lower corner on a value's coordinates (faces that touch)
  std::vector<float> qb(nq * 6);
  for (int64_t q = 0; q < nq; ++q)
  {
    const float h = r * (0.5f + 0.5f * (float)(q % 3));
    for (int d = 0; d < 3; ++d)
    {
      qb[6 * q + d] = sph[4 * q + d] - h;
      qb[6 * q + 3 + d] = sph[4 * q + d] + h;
    }
    if (q % 5 == 2)
      for (int d = 0; d < 3; ++d)
        qb[6 * q + d] = v[vs * ((q * 7) % n) + 3 * (geometry == 1) + d];
  }
  const int64_t bnnz = orc_query_spatial(t, 1, qb.data(), nq, off.data(), nullptr, nullptr,
                                         nullptr);
  std::vector<int32_t> bref(bnnz > 0 ? bnnz : 1);
  orc_query_spatial(t, 1, qb.data(), nq, off.data(), bref.data(), nullptr, nullptr);
  bool ok_box = true;
  err = 0;
  for (int64_t q = 0; q < nq && ok_box; ++q)
  {
    const float *c = &qb[6 * q];
    const axb::SpatialPred pr = axb::box_pred(axb::Box6{{c[0], c[1], c[2]}, {c[3], c[4], c[5]}});
    std::vector<int32_t> row;
    RowSink sink{&row};
    uint32_t vi = 0, vl = 0;
    if (geometry == 0)
    {
      if (node_format == 0)
        axb::rope_walk<axb::NODES_FP32, axb::GEOM_POINTS, axb::PRED_BOXES>(
            tv, pr, 0u, (uint32_t)total, ~0u, sink, vi, vl);
      else
        axb::rope_walk<axb::NODES_Q16, axb::GEOM_POINTS, axb::PRED_BOXES>(
            tv, pr, 0u, (uint32_t)total, ~0u, sink, vi, vl);
    }
    else
    {
      if (node_format == 0)
        axb::rope_walk<axb::NODES_FP32, axb::GEOM_BOXES, axb::PRED_BOXES>(
            tv, pr, 0u, (uint32_t)total, ~0u, sink, vi, vl);
      else
        axb::rope_walk<axb::NODES_Q16, axb::GEOM_BOXES, axb::PRED_BOXES>(
            tv, pr, 0u, (uint32_t)total, ~0u, sink, vi, vl);
    }
    ok_box = row.size() == (size_t)(off[q + 1] - off[q]) &&
             std::equal(row.begin(), row.end(), bref.begin() + off[q]);
  }
  std::snprintf(what, sizeof what, "box rows: geometry %d n %lld format %d", geometry,
                (long long)n, node_format);
  CHECK(ok_box && err == 0 && bnnz > 0, what);
  if (geometry == 0)
  {
    std::snprintf(what, sizeof what, "skip/stop forms: n %lld format %d", (long long)n,
                  node_format);
    CHECK(ok_skip, what);
  }
  orc_free(t);
}

// Stops after `limit` hits.
struct LimitSink
{
  std::vector<int32_t> *row;
  size_t limit;
  bool hit(int32_t v)
  {
    row->push_back(v);
    return row->size() < limit;
  }
};

// Rays against n boxes: the rope walk with ray predicates against the
// oracle's brute force (sets, tree order), and the best-first walk against
// the oracle's (distance, index)-sorted rows, whole and cut short.
static void run_rays(int64_t n, unsigned seed, int node_format, float size = 0.1f)
{
  std::mt19937 gen(seed);
  std::uniform_real_distribution<float> U(0.f, 1.f), O(-0.2f, 1.2f);
  std::normal_distribution<float> Nd(0.f, 1.f);
  std::vector<float> v(n * 6);
  for (int64_t i = 0; i < n; ++i)
  {
    const float x = U(gen), y = U(gen), z = U(gen);
    const float e = 0.004f + size * U(gen) * U(gen);
    v[6 * i] = x, v[6 * i + 1] = y, v[6 * i + 2] = z;
    v[6 * i + 3] = x + e, v[6 * i + 4] = y + e * U(gen), v[6 * i + 5] = z + e;
  }
  const int64_t nq = 600;
  std::vector<float> o(3 * nq), d(3 * nq), rays(6 * nq);
  for (int64_t q = 0; q < nq; ++q)
  {
    for (int a = 0; a < 3; ++a)
      o[3 * q + a] = O(gen), d[3 * q + a] = Nd(gen);
    if (q % 5 == 1) // along an axis
    {
      const int a = (int)(q % 3);
      const float keep = d[3 * q + a];
      d[3 * q] = d[3 * q + 1] = d[3 * q + 2] = 0.f;
      d[3 * q + a] = keep;
    }
    if (q % 7 == 3 && n > 0) // origin on a box face
    {
      const int64_t j = (int64_t)(gen() % (uint32_t)n);
      for (int a = 0; a < 3; ++a)
        o[3 * q + a] = 0.5f * (v[6 * j + a] + v[6 * j + 3 + a]);
      o[3 * q + q % 3] = v[6 * j + q % 3];
    }
  }
  orc_make_rays(o.data(), d.data(), nq, rays.data());
  void *t = orc_build(v.data(), n, 1);
  std::vector<uint32_t> words((size_t)(n * 4 + 16) * 4, 0u);
  const int64_t total = orc_linearize(t, node_format, AXB_FLAT_SUBTREE_MAX, words.data());
  std::vector<uint4> slots((size_t)total + 8, make_uint4(0u, 0u, 0u, 0u));
  std::memcpy(slots.data(), words.data(), (size_t)total * sizeof(uint4));
  HostTree tv;
  tv.slots = slots.data();
  float b[6];
  orc_bounds(t, b);
  for (int a = 0; a < 3; ++a)
  {
    tv.scene.lo[a] = b[a], tv.scene.hi[a] = b[3 + a];
    tv.ss[a] = axb::q16_scale(b[a], b[3 + a]);
  }
  uint32_t err = 0;
  tv.err = &err;
  std::vector<int64_t> off(nq + 1), soff(nq + 1);
  const int64_t nnz = orc_bruteforce_rays_ordered(v.data(), n, rays.data(), nq, -1, off.data(),
                                                  nullptr);
  std::vector<int32_t> ref(nnz > 0 ? nnz : 1), sref(nnz > 0 ? nnz : 1);
  orc_bruteforce_rays_ordered(v.data(), n, rays.data(), nq, -1, off.data(), ref.data());
  orc_bruteforce_spatial(v.data(), n, 1, 3, rays.data(), nq, soff.data(), sref.data());
  bool ok_spatial = true, ok_ordered = true, ok_cut = true, ok_sorted = true;
  int overflows = 0;
  for (int64_t q = 0; q < nq; ++q)
  {
    const float *r6 = rays.data() + 6 * q;
    const axb::SpatialPred pr = axb::ray_spatial_pred(r6);
    std::vector<int32_t> row, ord, cut;
    RowSink sink{&row}, so{&ord};
    uint32_t vi = 0, vl = 0;
    if (node_format == 0)
      axb::rope_walk<axb::NODES_FP32, axb::GEOM_BOXES, axb::PRED_RAYS>(
          tv, pr, 0u, (uint32_t)total, ~0u, sink, vi, vl);
    else
      axb::rope_walk<axb::NODES_Q16, axb::GEOM_BOXES, axb::PRED_RAYS>(
          tv, pr, 0u, (uint32_t)total, ~0u, sink, vi, vl);
    std::sort(row.begin(), row.end());
    ok_spatial = ok_spatial && row.size() == (size_t)(soff[q + 1] - soff[q]) &&
                 std::equal(row.begin(), row.end(), sref.begin() + soff[q]);
    // slots of the hits, sorted into the ordered row
    std::vector<int32_t> srow;
    RowSink ss{&srow};
    if (node_format == 0)
      axb::rope_walk<axb::NODES_FP32, axb::GEOM_BOXES, axb::PRED_RAYS, true>(
          tv, pr, 0u, (uint32_t)total, ~0u, ss, vi, vl);
    else
      axb::rope_walk<axb::NODES_Q16, axb::GEOM_BOXES, axb::PRED_RAYS, true>(
          tv, pr, 0u, (uint32_t)total, ~0u, ss, vi, vl);
    axb::sort_ray_row(tv, pr.ray, srow.data(), (int64_t)srow.size());
    ok_sorted = ok_sorted && srow.size() == (size_t)(off[q + 1] - off[q]) &&
                std::equal(srow.begin(), srow.end(), ref.begin() + off[q]);
    const size_t k = 1 + (size_t)(q % 4);
    LimitSink sc{&cut, k};
    bool fine;
    if (node_format == 0)
      fine = axb::ordered_walk<axb::NODES_FP32>(tv, pr.ray, so, vi, vl) &&
             axb::ordered_walk<axb::NODES_FP32>(tv, pr.ray, sc, vi, vl);
    else
      fine = axb::ordered_walk<axb::NODES_Q16>(tv, pr.ray, so, vi, vl) &&
             axb::ordered_walk<axb::NODES_Q16>(tv, pr.ray, sc, vi, vl);
    const size_t m = (size_t)(off[q + 1] - off[q]);
    if (!fine)
    {
      // a heap of 64 pending entries was not enough: reported, never wrong
      ++overflows;
      ok_ordered = ok_ordered && (err & axb::KERR_STACK) != 0;
      err &= ~axb::KERR_STACK;
      continue;
    }
    ok_ordered = ok_ordered && ord.size() == m &&
                 std::equal(ord.begin(), ord.end(), ref.begin() + off[q]);
    ok_cut = ok_cut && cut.size() == std::min(k, m) &&
             std::equal(cut.begin(), cut.end(), ref.begin() + off[q]);
  }
  char what[160];
  std::snprintf(what, sizeof what, "ray rows (rope walk): n %lld format %d", (long long)n,
                node_format);
  CHECK(ok_spatial && err == 0, what);
  std::snprintf(what, sizeof what, "ordered rows (best-first): n %lld format %d, %d heap overflows",
                (long long)n, node_format, overflows);
  CHECK(ok_ordered && err == 0, what);
  if (overflows)
    std::printf("note: %s\n", what);
  std::snprintf(what, sizeof what, "ordered rows cut at k: n %lld format %d", (long long)n,
                node_format);
  CHECK(ok_cut, what);
  std::snprintf(what, sizeof what, "ordered rows by slot sort: n %lld format %d", (long long)n,
                node_format);
  CHECK(ok_sorted, what);
  std::snprintf(what, sizeof what, "ray cases hit boxes: n %lld nnz %lld", (long long)n,
                (long long)nnz);
  CHECK(n < 1000 || nnz > nq, what);
  orc_free(t);
}

int main()
{
  for (int fmt : {0, 1})
  {
    for (int64_t n : {1, 2, 3, 5, 9, 40, 700, 6000, 60000})
      run_rays(n, 23u + (unsigned)n, fmt);
    run_rays(30000, 5u, fmt, 0.4f); // rows of many hundreds
  }
  for (int fmt : {0, 1})
  {
    for (int64_t n : {1, 2, 3, 4, 5, 6, 9, 17, 100, 1000, 20000})
      run_case(0, n, 7u + (unsigned)n, n < 100 ? 0.5f : 0.05f, fmt);
    for (int64_t n : {2, 5, 300, 5000})
      run_case(1, n, 11u + (unsigned)n, 0.03f, fmt);
  }
  std::printf("%d passed, %d failed\n", g_pass, g_fail);
  return g_fail == 0 ? 0 : 1;
}
