// test_callbacks.hip -- the header-only callback traversals
// (include/ArborX_Callbacks.hpp through include/ArborX.hpp) exercised like
// the reference's test/tstQueryTreeCallbacks.cpp on the GPU:
//   callback_spatial_predicate / callback_nearest_predicate (:95-179):
//     inline callbacks inserting {index, distance to origin};
//   callback_early_exit (:181-229): counts(i)++ < i, else early_exit;
//   callback_with_attachment_* (:264-380): data added to each insert;
// plus void callbacks and query(PerThread, ...) in a user kernel against the
// CRS query on random points, for both node formats; the reference cases,
// early exit and the CRS-through-callbacks check also over the brute-force
// index (the reference runs its callback suite over BruteForce too,
// test/CMakeLists.txt:87-124).
#include <ArborX.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <set>
#include <utility>
#include <vector>

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

using namespace ArborX;

struct PairIndexDistance
{
  int index;
  float distance;
  bool operator<(PairIndexDistance const &o) const
  {
    return index < o.index || (index == o.index && distance < o.distance);
  }
  bool operator==(PairIndexDistance const &o) const
  {
    return index == o.index && distance == o.distance;
  }
};

// CustomInlineCallback(WithAttachment) of tstQueryTreeCallbacks.cpp:27-42,
// 231-247: the attached value is looked up by the predicate's position.
struct InlineCallback
{
  Point const *points;
  float const *attached; // may be null
  template <class Predicate, class Insert>
  __device__ void operator()(Predicate const &p, int index, Insert const &insert) const
  {
    Point const origin{{0.f, 0.f, 0.f}};
    float d = distance(points[index], origin);
    float data = attached ? attached[getData(p)] : 0.f;
    insert(PairIndexDistance{index, data + d});
  }
};

static std::vector<std::multiset<PairIndexDistance>>
rows_of(DeviceVector<PairIndexDistance> const &v, DeviceVector<int64_t> const &o)
{
  auto hv = v.to_host();
  auto ho = o.to_host();
  std::vector<std::multiset<PairIndexDistance>> r;
  for (size_t q = 0; q + 1 < ho.size(); ++q)
    r.emplace_back(hv.begin() + ho[q], hv.begin() + ho[q + 1]);
  return r;
}

// Index factories: the tree in either node format, or the brute-force index.
struct MakeBVH
{
  NodeFormat fmt;
  template <class V>
  BoundingVolumeHierarchy operator()(HipStream const &space, DeviceVector<V> const &v) const
  {
    return BoundingVolumeHierarchy(space, v, fmt);
  }
};
struct MakeBruteForce
{
  template <class V>
  BruteForce operator()(HipStream const &space, DeviceVector<V> const &v) const
  {
    return BruteForce(space, v);
  }
};

template <class Make>
static void reference_cases(HipStream const &space, Make make)
{
  int const n = 10;
  std::vector<Point> hp(n);
  for (int i = 0; i < n; ++i)
    hp[i] = {{(float)i, (float)i, (float)i}};
  DeviceVector<Point> points(hp);
  auto tree = make(space, points);

  for (float delta : {0.f, 5.f})
  {
    std::multiset<PairIndexDistance> want;
    for (int i = 0; i < n; ++i)
    {
      Point const origin{{0.f, 0.f, 0.f}};
      want.insert({i, delta + distance(hp[i], origin)});
    }
    DeviceVector<float> att(std::vector<float>{delta});
    InlineCallback cb{points.data(), att.data()};

    DeviceVector<Box> qb(std::vector<Box>{tree.bounds()});
    DeviceVector<PairIndexDistance> out;
    DeviceVector<int64_t> off;
    tree.query(space, intersects(qb), cb, out, off);
    auto r = rows_of(out, off);
    CHECK(r.size() == 1 && r[0] == want, "callback_spatial_predicate");

    DeviceVector<Point> origin(std::vector<Point>{{{0.f, 0.f, 0.f}}});
    tree.query(space, nearest(origin, n), cb, out, off);
    r = rows_of(out, off);
    CHECK(r.size() == 1 && r[0] == want, "callback_nearest_predicate");
    // nearest callbacks arrive in (distance, index) order
    auto hv = out.to_host();
    bool sorted = true;
    for (int i = 0; i < n; ++i)
      sorted = sorted && hv[i].index == i;
    CHECK(sorted, "nearest callbacks in distance order");
  }
}

// Experimental_CustomCallbackEarlyExit (tstQueryTreeCallbacks.cpp:181-229)
struct EarlyExit
{
  int *counts;
  template <class Predicate>
  __device__ CallbackTreeTraversalControl operator()(Predicate const &p, int) const
  {
    int i = getData(p);
    if (atomicAdd(&counts[i], 1) < i)
      return CallbackTreeTraversalControl::normal_continuation;
    return CallbackTreeTraversalControl::early_exit;
  }
};

template <class Make>
static void early_exit(HipStream const &space, Make make)
{
  std::vector<Box> hb;
  for (int i = 0; i < 4; ++i)
  {
    float f = (float)i;
    hb.push_back(Box({{f, f, f}}, {{f, f, f}}));
  }
  DeviceVector<Box> boxes(hb);
  auto tree = make(space, boxes);
  Box b = tree.bounds();
  DeviceVector<Box> q(std::vector<Box>{b, b, b, b});
  DeviceVector<int> counts(std::vector<int>{0, 0, 0, 0});
  tree.query(space, intersects(q), EarlyExit{counts.data()});
  CHECK((counts.to_host() == std::vector<int>{1, 2, 3, 4}), "callback_early_exit");
}

struct CountHits
{
  int *counts;
  template <class Predicate>
  __device__ void operator()(Predicate const &p, int) const
  {
    atomicAdd(&counts[getData(p)], 1);
  }
};

// query(PerThread, ...) from a user kernel: counts and index sums per query.
__global__ void per_thread_kernel(Details::DeviceTree tree, Sphere const *spheres, int nq,
                                  int *counts, long long *sums)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq)
    return;
  int c = 0;
  long long s = 0;
  tree.query(Experimental::PerThread{}, Intersects<Sphere>{spheres[i], i},
             [&](Intersects<Sphere> const &, int v) {
               ++c;
               s += v;
             });
  counts[i] = c;
  sums[i] = s;
}

static void random_points(HipStream const &space, NodeFormat fmt)
{
  int const n = 20000, nq = 5000;
  std::mt19937 gen(7);
  std::uniform_real_distribution<float> U(0.f, 1.f);
  std::vector<Point> hp(n);
  for (auto &p : hp)
    p = {{U(gen), U(gen), U(gen)}};
  std::vector<Sphere> hs(nq);
  for (auto &s : hs)
    s = Sphere{{{U(gen), U(gen), U(gen)}}, 0.03f * U(gen)};
  DeviceVector<Point> pts(hp);
  DeviceVector<Sphere> sph(hs);
  BoundingVolumeHierarchy tree(space, pts, fmt);

  DeviceVector<int> I;
  DeviceVector<int64_t> O;
  tree.query(space, intersects(sph), I, O);
  auto ho = O.to_host();
  auto hi = I.to_host();

  for (bool sorted : {true, false})
  {
    DeviceVector<int> counts(std::vector<int>(nq, 0));
    tree.query(space, intersects(sph), CountHits{counts.data()},
               Experimental::TraversalPolicy().setPredicateSorting(sorted));
    auto hc = counts.to_host();
    bool ok = true;
    for (int q = 0; q < nq; ++q)
      ok = ok && hc[q] == ho[q + 1] - ho[q];
    CHECK(ok, "void callback counts equal the CRS row lengths");
  }

  DeviceVector<int> c2(nq);
  DeviceVector<long long> s2(nq);
  per_thread_kernel<<<(nq + 255) / 256, 256, 0, space.stream>>>(tree.device_view(), sph.data(),
                                                                nq, c2.data(), s2.data());
  Details::hip_check(hipStreamSynchronize(space.stream), "per_thread_kernel");
  auto hc = c2.to_host();
  auto hsum = s2.to_host();
  bool ok = true;
  for (int q = 0; q < nq; ++q)
  {
    long long s = 0;
    for (int64_t j = ho[q]; j < ho[q + 1]; ++j)
      s += hi[j];
    ok = ok && hc[q] == ho[q + 1] - ho[q] && hsum[q] == s;
  }
  CHECK(ok, "query(PerThread) equals the CRS query");

  // CRS through an inserting callback equals the library's CRS, row order
  // included (same rope walk)
  DeviceVector<int> vals;
  DeviceVector<int64_t> off2;
  tree.query(
      space, intersects(sph),
      [] __device__(Intersects<Sphere> const &, int v, auto const &out) { out(v); }, vals, off2);
  CHECK(off2.to_host() == ho && vals.to_host() == hi, "default-callback CRS equals query");
}

// First hit of each ray: the ordered callback stops at once.
struct FirstHit
{
  int *first;
  __device__ CallbackTreeTraversalControl operator()(
      OrderedSpatial<Experimental::Ray> const &p, int v) const
  {
    first[getData(p)] = v;
    return CallbackTreeTraversalControl::early_exit;
  }
};

// query(PerThread, ordered ray, callback) from a user kernel: the first two
// hits of each ray.
__global__ void per_thread_rays(Details::DeviceTree tree, Experimental::Ray const *rays, int nq,
                                int *firsts)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq)
    return;
  int c = 0;
  firsts[2 * i] = firsts[2 * i + 1] = -1;
  query(Experimental::PerThread{}, tree, OrderedSpatial<Experimental::Ray>{rays[i], i},
        [&](OrderedSpatial<Experimental::Ray> const &, int v) {
          firsts[2 * i + c] = v;
          return ++c < 2 ? CallbackTreeTraversalControl::normal_continuation
                         : CallbackTreeTraversalControl::early_exit;
        });
}

// Ray callbacks over random boxes against the library's CRS rows: the
// ordered callback's first hit and first two hits (PerThread), the CRS
// built through an inserting ordered callback (same order), the void
// Intersects<Ray> callback counts, and nearest(ray, k) per-value callbacks.
static void random_rays(HipStream const &space, NodeFormat fmt)
{
  using Experimental::Ray;
  int const n = 20000, nq = 4000;
  std::mt19937 gen(11);
  std::uniform_real_distribution<float> U(0.f, 1.f), O(-0.2f, 1.2f);
  std::normal_distribution<float> Nd(0.f, 1.f);
  std::vector<Box> hb(n);
  for (auto &b : hb)
  {
    const float x = U(gen), y = U(gen), z = U(gen), e = 0.002f + 0.03f * U(gen);
    b = Box({{x, y, z}}, {{x + e, y + e, z + e}});
  }
  std::vector<Ray> hr(nq);
  for (int q = 0; q < nq; ++q)
  {
    Point dir{{Nd(gen), Nd(gen), Nd(gen)}};
    if (q % 5 == 1)
      dir[(q / 5) % 3] = 0.f;
    hr[q] = Ray({{O(gen), O(gen), O(gen)}}, dir);
  }
  DeviceVector<Box> boxes(hb);
  DeviceVector<Ray> rays(hr);
  BoundingVolumeHierarchy tree(space, boxes, fmt);
  DeviceVector<int> I;
  DeviceVector<int64_t> Off;
  tree.query(space, Experimental::ordered_intersects(rays), I, Off);
  auto ho = Off.to_host();
  auto hi = I.to_host();
  CHECK(ho.back() > nq, "rays hit boxes");

  DeviceVector<int> first(std::vector<int>(nq, -1));
  tree.query(space, Experimental::ordered_intersects(rays), FirstHit{first.data()});
  auto hf = first.to_host();
  DeviceVector<int> two(2 * nq);
  per_thread_rays<<<(nq + 255) / 256, 256, 0, space.stream>>>(tree.device_view(), rays.data(),
                                                              nq, two.data());
  Details::hip_check(hipStreamSynchronize(space.stream), "per_thread_rays");
  auto h2 = two.to_host();
  bool ok1 = true, ok2 = true;
  for (int q = 0; q < nq; ++q)
  {
    const int64_t m = ho[q + 1] - ho[q];
    ok1 = ok1 && hf[q] == (m > 0 ? hi[ho[q]] : -1);
    ok2 = ok2 && h2[2 * q] == (m > 0 ? hi[ho[q]] : -1) && h2[2 * q + 1] == (m > 1 ? hi[ho[q] + 1] : -1);
  }
  CHECK(ok1, "ordered callback with early exit: the first hit");
  CHECK(ok2, "query(PerThread) ordered: the first two hits");

  DeviceVector<int> vals;
  DeviceVector<int64_t> off2;
  tree.query(
      space, Experimental::ordered_intersects(rays),
      [] __device__(OrderedSpatial<Ray> const &, int v, auto const &out) { out(v); }, vals, off2);
  CHECK(off2.to_host() == ho && vals.to_host() == hi, "ordered inserting callback equals CRS");

  DeviceVector<int> counts(std::vector<int>(nq, 0));
  tree.query(space, intersects(rays), CountHits{counts.data()});
  auto hc = counts.to_host();
  bool ok = true;
  for (int q = 0; q < nq; ++q)
    ok = ok && hc[q] == ho[q + 1] - ho[q];
  CHECK(ok, "Intersects<Ray> callback counts equal the ordered row lengths");
  if (fmt == NodeFormat::FP32)
  {
    BruteForce bf(space, boxes);
    DeviceVector<int> bcounts(std::vector<int>(nq, 0));
    bf.query(space, intersects(rays), CountHits{bcounts.data()});
    CHECK(bcounts.to_host() == hc, "brute force: Intersects<Ray> callback counts");
  }

  DeviceVector<int> NI;
  DeviceVector<int64_t> NO;
  tree.query(space, nearest(rays, 3), NI, NO);
  DeviceVector<int> nvals;
  DeviceVector<int64_t> noff;
  tree.query(
      space, nearest(rays, 3),
      [] __device__(Nearest<Ray> const &, int v, auto const &out) { out(v); }, nvals, noff);
  CHECK(noff.to_host() == NO.to_host() && nvals.to_host() == NI.to_host(),
        "nearest(ray) inserting callback equals its CRS");
  auto nho = NO.to_host();
  auto nhi = NI.to_host();
  ok = true;
  for (int q = 0; q < nq && ok; ++q)
  {
    const int64_t m = std::min<int64_t>(3, ho[q + 1] - ho[q]);
    ok = nho[q + 1] - nho[q] == m && std::equal(nhi.begin() + nho[q], nhi.begin() + nho[q + 1],
                                                hi.begin() + ho[q]);
  }
  CHECK(ok, "nearest(ray, 3) rows are the first 3 ordered hits");

  // rays need boxes
  DeviceVector<Point> pts(std::vector<Point>{{{0.f, 0.f, 0.f}}, {{1.f, 1.f, 1.f}}});
  BoundingVolumeHierarchy ptree(space, pts, fmt);
  bool threw = false;
  try
  {
    ptree.query(space, intersects(rays), CountHits{counts.data()});
  }
  catch (std::invalid_argument const &)
  {
    threw = true;
  }
  CHECK(threw, "ray callbacks over points throw");
}

// The brute-force index's callbacks against its own CRS query: counts per
// predicate, and the CRS built through an inserting callback equals it,
// row order included (values in index order on both sides).
static void bruteforce_random(HipStream const &space)
{
  int const n = 5000, nq = 2000;
  std::mt19937 gen(9);
  std::uniform_real_distribution<float> U(0.f, 1.f);
  std::vector<Box> hb(n);
  for (auto &b : hb)
  {
    float x = U(gen), y = U(gen), z = U(gen), e = 0.02f * U(gen);
    b = Box({{x, y, z}}, {{x + e, y + e, z + e}});
  }
  std::vector<Sphere> hs(nq);
  for (auto &s : hs)
    s = Sphere{{{U(gen), U(gen), U(gen)}}, 0.05f * U(gen)};
  DeviceVector<Box> boxes(hb);
  DeviceVector<Sphere> sph(hs);
  BruteForce bf(space, boxes);
  DeviceVector<int> I;
  DeviceVector<int64_t> O;
  bf.query(space, intersects(sph), I, O);
  auto ho = O.to_host();
  auto hi = I.to_host();
  DeviceVector<int> counts(std::vector<int>(nq, 0));
  bf.query(space, intersects(sph), CountHits{counts.data()});
  auto hc = counts.to_host();
  bool ok = ho.back() > 0;
  for (int q = 0; q < nq; ++q)
    ok = ok && hc[q] == ho[q + 1] - ho[q];
  CHECK(ok, "brute force: void callback counts equal the CRS row lengths");
  DeviceVector<int> vals;
  DeviceVector<int64_t> off2;
  bf.query(
      space, intersects(sph),
      [] __device__(Intersects<Sphere> const &, int v, auto const &out) { out(v); }, vals, off2);
  CHECK(off2.to_host() == ho && vals.to_host() == hi,
        "brute force: default-callback CRS equals query");
}

// The XCD block remap of the traversal kernels must be a bijection of
// [0, nblocks) (host-side check of the device helper).
// tstQueryTreeCallbackQueryPerThread.cpp:36-110: ten points (i, i, i);
// thread i queries the box [i - .5, i + .5]^3 with query(PerThread, ...):
// intersects(box) -- every match is point i -- and ordered_intersects(box)
// with a callback that records whether its first value is i and returns
// early_exit.
__global__ void per_thread_reference(Details::DeviceTree tree, int n, int *ok)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const float c = (float)i;
  const Box box({{c - .5f, c - .5f, c - .5f}}, {{c + .5f, c + .5f, c + .5f}});
  int s1 = 1, s2 = 0, calls = 0;
  query(Experimental::PerThread{}, tree, Intersects<Box>{box, i},
        [&](Intersects<Box> const &, int v) { s1 = s1 && v == i; });
  query(Experimental::PerThread{}, tree, OrderedSpatial<Box>{box, i},
        [&](OrderedSpatial<Box> const &, int v) {
          s2 = v == i;
          ++calls;
          return CallbackTreeTraversalControl::early_exit;
        });
  ok[i] = s1 && s2 && calls == 1;
}
static void per_thread_reference_case(HipStream const &space, NodeFormat fmt)
{
  const int n = 10;
  std::vector<Point> hp(n);
  for (int i = 0; i < n; ++i)
    hp[i] = Point{{(float)i, (float)i, (float)i}};
  DeviceVector<Point> points(hp);
  BoundingVolumeHierarchy tree(space, points, fmt);
  DeviceVector<int> ok(std::vector<int>(n, 0));
  per_thread_reference<<<1, 64, 0, space.stream>>>(tree.device_view(), n, ok.data());
  Details::hip_check(hipStreamSynchronize(space.stream), "per_thread_reference");
  CHECK(ok.to_host() == std::vector<int>(n, 1), "PerThread intersects / ordered_intersects(box)");
}

// Ordered spheres, boxes and points over random points and boxes: the
// callback form (every value, in order; early exit after the first three)
// against the CRS rows of the same predicates.
template <class G>
struct OrderedPrefix
{
  int *out; // 3 per predicate
  __device__ CallbackTreeTraversalControl operator()(OrderedSpatial<G> const &p, int v) const
  {
    int *o = out + 3 * getData(p);
    int j = 0;
    while (j < 3 && o[j] >= 0)
      ++j;
    if (j < 3)
      o[j] = v;
    return j + 1 < 3 ? CallbackTreeTraversalControl::normal_continuation
                     : CallbackTreeTraversalControl::early_exit;
  }
};
template <class G, class Batch, class Tree>
static void ordered_prefix_check(HipStream const &space, Tree const &tree, Batch const &b,
                                 int nq, int n, const char *what)
{
  DeviceVector<int> I;
  DeviceVector<int64_t> Off;
  tree.query(space, b, I, Off);
  auto ho = Off.to_host();
  auto hi = I.to_host();
  bool ok = ho.size() == (size_t)nq + 1 && ho.back() == (int64_t)nq * n;
  DeviceVector<int> pre(std::vector<int>(3 * nq, -1));
  tree.query(space, b, OrderedPrefix<G>{pre.data()});
  auto hpre = pre.to_host();
  for (int q = 0; ok && q < nq; ++q)
    for (int j = 0; j < 3; ++j)
      ok = ok && hpre[3 * q + j] == (j < n ? hi[ho[q] + j] : -1);
  // every value once per row
  for (int q = 0; ok && q < nq; ++q)
  {
    std::vector<int> r(hi.begin() + ho[q], hi.begin() + ho[q + 1]);
    std::sort(r.begin(), r.end());
    for (int j = 0; j < n; ++j)
      ok = ok && r[j] == j;
  }
  CHECK(ok, what);
}
static void random_ordered(HipStream const &space, NodeFormat fmt)
{
  const int n = 50, nq = 300; // the ordered walk's heap holds the few pending entries
  std::mt19937 gen(21);
  std::uniform_real_distribution<float> U(0.f, 1.f), O(-0.2f, 1.2f);
  std::vector<Point> hp(n);
  std::vector<Box> hb(n);
  for (int i = 0; i < n; ++i)
  {
    hp[i] = Point{{U(gen), U(gen), U(gen)}};
    const float e = 0.05f * U(gen);
    hb[i] = Box(hp[i], Point{{hp[i][0] + e, hp[i][1] + e, hp[i][2] + e}});
  }
  std::vector<Sphere> hs(nq);
  std::vector<Box> hq(nq);
  std::vector<Point> hc(nq);
  for (int q = 0; q < nq; ++q)
  {
    hc[q] = Point{{O(gen), O(gen), O(gen)}};
    hs[q] = Sphere{hc[q], 0.2f * U(gen)};
    const float e = 0.2f * U(gen);
    hq[q] = Box(hc[q], Point{{hc[q][0] + e, hc[q][1] + e, hc[q][2] + e}});
  }
  DeviceVector<Point> points(hp), centres(hc);
  DeviceVector<Box> boxes(hb), qboxes(hq);
  DeviceVector<Sphere> spheres(hs);
  BoundingVolumeHierarchy tp(space, points, fmt), tb(space, boxes, fmt);
  ordered_prefix_check<Sphere>(space, tp, Experimental::ordered_intersects(spheres), nq, n,
                               "ordered spheres over points: callback prefix = CRS");
  ordered_prefix_check<Box>(space, tp, Experimental::ordered_intersects(qboxes), nq, n,
                            "ordered boxes over points: callback prefix = CRS");
  ordered_prefix_check<Point>(space, tp, Experimental::ordered_intersects(centres), nq, n,
                              "ordered points over points: callback prefix = CRS");
  ordered_prefix_check<Sphere>(space, tb, Experimental::ordered_intersects(spheres), nq, n,
                               "ordered spheres over boxes: callback prefix = CRS");
  ordered_prefix_check<Box>(space, tb, Experimental::ordered_intersects(qboxes), nq, n,
                            "ordered boxes over boxes: callback prefix = CRS");
  ordered_prefix_check<Point>(space, tb, Experimental::ordered_intersects(centres), nq, n,
                              "ordered points over boxes: callback prefix = CRS");
}

static void xcd_remap()
{
  bool ok = true;
  for (uint32_t nb : {1u, 2u, 7u, 8u, 9u, 15u, 16u, 17u, 100u, 1001u, 39063u})
  {
    std::vector<int> seen(nb, 0);
    for (uint32_t b = 0; b < nb; ++b)
    {
      uint32_t l = axb::xcd_block(b, nb);
      if (l >= nb)
        ok = false;
      else
        ++seen[l];
    }
    for (int s : seen)
      ok = ok && s == 1;
    // every XCD (b % 8) gets one contiguous range
    for (uint32_t b = 8; b < nb; ++b)
      ok = ok && axb::xcd_block(b, nb) == axb::xcd_block(b - 8, nb) + 1;
  }
  CHECK(ok, "xcd_block is a bijection with contiguous ranges per XCD");
}

int main()
{
  HipStream space;
  xcd_remap();
  for (NodeFormat fmt : {NodeFormat::FP32, NodeFormat::Q16})
  {
    reference_cases(space, MakeBVH{fmt});
    early_exit(space, MakeBVH{fmt});
    random_points(space, fmt);
    random_rays(space, fmt);
    per_thread_reference_case(space, fmt);
    random_ordered(space, fmt);
  }
  reference_cases(space, MakeBruteForce{});
  early_exit(space, MakeBruteForce{});
  bruteforce_random(space);
  // empty tree: callbacks are never called
  BoundingVolumeHierarchy empty;
  DeviceVector<Box> q(std::vector<Box>{Box({{0.f, 0.f, 0.f}}, {{1.f, 1.f, 1.f}})});
  DeviceVector<int> counts(std::vector<int>{0});
  empty.query(space, intersects(q), CountHits{counts.data()});
  CHECK(counts.to_host()[0] == 0, "empty tree: no callback");
  std::printf("%d passed, %d failed\n", g_pass, g_fail);
  return g_fail == 0 ? 0 : 1;
}
