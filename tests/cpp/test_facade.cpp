// test_facade.cpp -- the C++ facade (include/ArborX.hpp) exercised like the
// reference's own Boost.Test suites, on the GPU:
//   tstQueryTreeDegenerate.cpp (empty / one leaf / two leaves / duplicates /
//   4096-leaf capacity), tstQueryTreeCallbacks.cpp:138-179 (collinear kNN),
//   tstQueryTreeManufacturedSolution.cpp (11^3 grid),
//   tstQueryTreeTraversalPolicy.cpp:101-130 (buffer policy, SearchException),
//   tstNeighborList.cpp:116-176 (full / half neighbour lists),
// plus a brute-force comparison on random points; the same suites over the
// brute-force index (ArborX::BruteForce).  Rows are compared as
// multisets (test/boost_ext/CompressedStorageComparison.hpp:21-61).
#include <ArborX.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <set>
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
using Rows = std::vector<std::multiset<int>>;

static Rows rows_of(DeviceVector<int> const &ind, DeviceVector<int64_t> const &off)
{
  auto i = ind.to_host();
  auto o = off.to_host();
  Rows r(o.empty() ? 0 : o.size() - 1);
  for (size_t q = 0; q + 1 < o.size(); ++q)
    r[q].insert(i.begin() + o[q], i.begin() + o[q + 1]);
  return r;
}
static Rows rows(std::vector<std::vector<int>> const &v)
{
  Rows r;
  for (auto &x : v)
    r.emplace_back(x.begin(), x.end());
  return r;
}
static Box pbox(float a) { return Box({{a, a, a}}, {{a, a, a}}); }

// Index factories: the same suites run over the BVH (both node formats)
// and the brute-force index, as the reference instantiates its query-tree
// suites for both (test/CMakeLists.txt:87-124).
struct MakeBVH
{
  using index_type = BoundingVolumeHierarchy;
  NodeFormat fmt;
  template <class V>
  index_type operator()(HipStream const &space, DeviceVector<V> const &v) const
  {
    return index_type(space, v, fmt);
  }
};
struct MakeBruteForce
{
  using index_type = BruteForce;
  template <class V>
  index_type operator()(HipStream const &space, DeviceVector<V> const &v) const
  {
    return index_type(space, v);
  }
};

template <class Make>
static void degenerate(HipStream const &space, Make make)
{
  // empty tree (tstQueryTreeDegenerate.cpp:24-120)
  {
    typename Make::index_type bvh;
    CHECK(bvh.empty() && bvh.size() == 0, "default tree is empty");
    DeviceVector<Box> none;
    auto bvh2 = make(space, none);
    CHECK(bvh2.empty(), "tree over no values is empty");
    Box b = bvh2.bounds();
    CHECK(b.minCorner()[0] == FLT_MAX && b.maxCorner()[0] == -FLT_MAX, "empty bounds");
    DeviceVector<Box> q(std::vector<Box>{Box(), Box()});
    DeviceVector<int> ind;
    DeviceVector<int64_t> off;
    bvh2.query(space, intersects(q), ind, off);
    CHECK(rows_of(ind, off) == rows({{}, {}}), "empty tree, box queries");
    DeviceVector<Point> p(std::vector<Point>{{{0, 0, 0}}, {{1, 1, 1}}});
    bvh2.query(space, nearest(p, 2), ind, off);
    CHECK(rows_of(ind, off) == rows({{}, {}}), "empty tree, nearest");
  }
  // single leaf (:123-225)
  {
    DeviceVector<Box> v(std::vector<Box>{Box({{0, 0, 0}}, {{1, 1, 1}})});
    auto bvh = make(space, v);
    CHECK(bvh.size() == 1, "one leaf");
    DeviceVector<Box> q(std::vector<Box>{pbox(5), pbox(.5f)});
    DeviceVector<int> ind;
    DeviceVector<int64_t> off;
    bvh.query(space, intersects(q), ind, off);
    CHECK(rows_of(ind, off) == rows({{}, {0}}), "one leaf, boxes");
    DeviceVector<Sphere> s(
        std::vector<Sphere>{{{{0, 0, 0}}, 1.f}, {{{1, 1, 1}}, 3.f}, {{{5, 5, 5}}, 2.f}});
    bvh.query(space, intersects(s), ind, off);
    CHECK(rows_of(ind, off) == rows({{0}, {0}, {}}), "one leaf, spheres");
    DeviceVector<Point> p(std::vector<Point>{{{0, 0, 0}}, {{4, 5, 1}}});
    bvh.query(space, nearest(p, 3), ind, off);
    CHECK(rows_of(ind, off) == rows({{0}, {0}}), "one leaf, nearest");
  }
  // two leaves (:227-331)
  {
    DeviceVector<Box> v(std::vector<Box>{pbox(0), pbox(1)});
    auto bvh = make(space, v);
    DeviceVector<Box> q(std::vector<Box>{Box(), Box({{0, 0, 0}}, {{1, 1, 1}}),
                                         Box({{.5f, .5f, .5f}}, {{1.5f, 1.5f, 1.5f}}), pbox(0)});
    DeviceVector<int> ind;
    DeviceVector<int64_t> off;
    bvh.query(space, intersects(q), ind, off);
    CHECK(rows_of(ind, off) == rows({{}, {0, 1}, {1}, {0}}), "two leaves, boxes");
    DeviceVector<Point> p(std::vector<Point>{{{0, 0, 0}}, {{1, 0, 0}}});
    bvh.query(space, nearest(p, 4), ind, off);
    CHECK(rows_of(ind, off) == rows({{0, 1}, {0, 1}}), "two leaves, nearest k > n");
  }
  // duplicated leaves (:333-368)
  {
    DeviceVector<Box> v(std::vector<Box>{pbox(0), pbox(1), pbox(1), pbox(1)});
    auto bvh = make(space, v);
    DeviceVector<Sphere> s(std::vector<Sphere>{
        {{{0, 0, 0}}, 1.f}, {{{1, 1, 1}}, 1.f}, {{{.5f, .5f, .5f}}, 1.f}});
    DeviceVector<int> ind;
    DeviceVector<int64_t> off;
    bvh.query(space, intersects(s), ind, off);
    CHECK(rows_of(ind, off) == rows({{0}, {1, 2, 3}, {0, 1, 2, 3}}), "duplicated leaves");
  }
  // stack capacity: 4096 diagonal boxes (:375-446)
  {
    std::vector<Box> b;
    for (int i = 0; i < 4096; ++i)
      b.push_back(Box({{(float)i, (float)i, (float)i}},
                      {{(float)i + 1, (float)i + 1, (float)i + 1}}));
    DeviceVector<Box> v(b);
    auto bvh = make(space, v);
    DeviceVector<Box> q(std::vector<Box>{Box(), Box({{0, 0, 0}}, {{4096, 4096, 4096}})});
    DeviceVector<int> ind;
    DeviceVector<int64_t> off;
    bvh.query(space, intersects(q), ind, off);
    CHECK(off.to_host().back() == 4096, "spatial query over 4096 leaves");
    DeviceVector<Point> p(std::vector<Point>{{{0, 0, 0}}});
    bvh.query(space, nearest(p, 4096), ind, off);
    CHECK(off.to_host().back() == 4096, "nearest k = n = 4096");
  }
}

template <class Make>
static void callbacks_nearest(HipStream const &space, Make make)
{
  // tstQueryTreeCallbacks.cpp:138-179: points (i,i,i), nearest(origin, 10)
  std::vector<Point> pts;
  for (int i = 0; i < 10; ++i)
    pts.push_back({{(float)i, (float)i, (float)i}});
  DeviceVector<Point> v(pts);
  auto bvh = make(space, v);
  DeviceVector<Point> q(std::vector<Point>{{{0, 0, 0}}});
  DeviceVector<int> ind;
  DeviceVector<int64_t> off;
  DeviceVector<float> dist;
  bvh.query(space, nearest(q, 10), ind, off, dist);
  auto i = ind.to_host();
  auto d = dist.to_host();
  bool ok = i.size() == 10;
  for (int j = 0; ok && j < 10; ++j)
  {
    float s = 0, t = (float)j;
    for (int a = 0; a < 3; ++a)
      s += t * t;
    ok = i[j] == j && d[j] == std::sqrt(s);
  }
  CHECK(ok, "collinear nearest: indices and distances");
}

template <class Make>
static void manufactured(HipStream const &space, Make make)
{
  // tstQueryTreeManufacturedSolution.cpp:29-270 (cases i and ii)
  const int nx = 11, ny = 11, nz = 11, n = nx * ny * nz;
  const float h = 100.f / 10.f;
  auto ind = [&](int i, int j, int k) { return i + j * nx + k * nx * ny; };
  std::vector<Box> boxes(n), q(n);
  std::vector<std::vector<int>> ref(n);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k)
      {
        Point p{{i * h, j * h, k * h}};
        boxes[ind(i, j, k)] = Box(p, p);
        q[ind(i, j, k)] = Box({{(i - 1) * h, (j - 1) * h, (k - 1) * h}},
                              {{(i + 1) * h, (j + 1) * h, (k + 1) * h}});
        for (int a = -1; a <= 1; ++a)
          for (int b = -1; b <= 1; ++b)
            for (int c = -1; c <= 1; ++c)
              if (i + a >= 0 && i + a < nx && j + b >= 0 && j + b < ny && k + c >= 0 &&
                  k + c < nz)
                ref[ind(i, j, k)].push_back(ind(i + a, j + b, k + c));
      }
  DeviceVector<Box> v(boxes);
  auto bvh = make(space, v);
  DeviceVector<int> I;
  DeviceVector<int64_t> O;
  bvh.query(space, intersects(v), I, O);
  std::vector<std::vector<int>> self(n);
  for (int i = 0; i < n; ++i)
    self[i] = {i};
  CHECK(rows_of(I, O) == rows(self), "manufactured (i): objects themselves");
  DeviceVector<Box> qd(q);
  bvh.query(space, intersects(qd), I, O);
  CHECK(rows_of(I, O) == rows(ref), "manufactured (ii): first neighbours");
}

template <class Make>
static void policy(HipStream const &space, Make make)
{
  // tstQueryTreeTraversalPolicy.cpp:101-130
  std::vector<Point> pts;
  for (int i = 0; i < 64; ++i)
    pts.push_back({{(float)(i % 4), (float)((i / 4) % 4), (float)(i / 16)}});
  DeviceVector<Point> v(pts);
  auto bvh = make(space, v);
  DeviceVector<int> I, I0;
  DeviceVector<int64_t> O, O0;
  bvh.query(space, Experimental::make_intersects(v, 1.5f), I0, O0);
  for (int b : {64, 2, 1})
  {
    bvh.query(space, Experimental::make_intersects(v, 1.5f), I, O,
              Experimental::TraversalPolicy().setBufferSize(b));
    CHECK(rows_of(I, O) == rows_of(I0, O0), "buffer policy gives the same rows");
  }
  bool threw = false;
  try
  {
    bvh.query(space, Experimental::make_intersects(v, 1.5f), I, O,
              Experimental::TraversalPolicy().setBufferSize(-1));
  }
  catch (SearchException const &)
  {
    threw = true;
  }
  CHECK(threw, "hard buffer overflow throws SearchException");
  bvh.query(space, Experimental::make_intersects(v, 1.5f), I, O,
            Experimental::TraversalPolicy().setPredicateSorting(false));
  CHECK(rows_of(I, O) == rows_of(I0, O0), "predicate sorting off gives the same rows");
}

template <class Make>
static void brute_force(HipStream const &space, Make make)
{
  std::mt19937 gen(42);
  std::uniform_real_distribution<float> U(0.f, 1.f);
  const int n = 20000, nq = 2000, k = 12;
  const float r = 0.06f;
  std::vector<Point> pts(n), qp(nq);
  for (auto &p : pts)
    p = {{U(gen), U(gen), U(gen)}};
  for (auto &p : qp)
    p = {{U(gen), U(gen), U(gen)}};
  DeviceVector<Point> v(pts), q(qp);
  auto bvh = make(space, v);
  DeviceVector<int> I;
  DeviceVector<int64_t> O;
  bvh.query(space, Experimental::make_intersects(q, r), I, O);
  Rows got = rows_of(I, O);
  DeviceVector<float> D;
  bvh.query(space, Experimental::make_nearest(q, k), I, O, D);
  auto ki = I.to_host();
  auto kd = D.to_host();
  bool ok_r = true, ok_k = true;
  for (int a = 0; a < nq; ++a)
  {
    std::multiset<int> want;
    std::vector<std::pair<float, int>> all;
    for (int b = 0; b < n; ++b)
    {
      float s = 0;
      for (int d = 0; d < 3; ++d)
      {
        float t = pts[b][d] - qp[a][d];
        s += t * t;
      }
      float dist = std::sqrt(s);
      if (dist <= r)
        want.insert(b);
      all.push_back({dist, b});
    }
    std::partial_sort(all.begin(), all.begin() + k, all.end());
    ok_r = ok_r && (got[a] == want);
    for (int j = 0; j < k; ++j)
      ok_k = ok_k && ki[a * k + j] == all[j].second && kd[a * k + j] == all[j].first;
  }
  CHECK(ok_r, "radius query equals brute force");
  CHECK(ok_k, "kNN equals brute force (distance, index)");
}

// Trees over attach_indices(values) (tstAttachIndices.cpp:19-29: unsigned
// indices by default, index = position): queries returning
// PairValueIndex{value, index} give the index form's rows with each value
// equal to values[index]; boxes keep their six floats; an empty tree gives
// empty rows; values of another type are refused.
static void attach_indices_values(HipStream const &space, NodeFormat fmt)
{
  static_assert(std::is_same_v<decltype(PairValueIndex<Point>{}.index), unsigned>);
  std::mt19937 gen(7);
  std::uniform_real_distribution<float> U(0.f, 1.f);
  const int n = 3000, nq = 500, k = 5;
  std::vector<Point> pts(n), qp(nq);
  for (auto &p : pts)
    p = {{U(gen), U(gen), U(gen)}};
  for (auto &p : qp)
    p = {{U(gen), U(gen), U(gen)}};
  DeviceVector<Point> v(pts), q(qp);
  BVH plain(space, v, fmt);
  BVH tree(space, Experimental::attach_indices(v), fmt);
  DeviceVector<int> I;
  DeviceVector<int64_t> O, O2;
  DeviceVector<PairValueIndex<Point>> P;
  bool ok = true;
  for (int form = 0; form < 2; ++form)
  {
    if (form == 0)
    {
      plain.query(space, Experimental::make_intersects(q, 0.08f), I, O);
      tree.query(space, Experimental::make_intersects(q, 0.08f), P, O2);
    }
    else
    {
      plain.query(space, Experimental::make_nearest(q, k), I, O);
      tree.query(space, Experimental::make_nearest(q, k), P, O2);
    }
    auto i = I.to_host();
    auto pv = P.to_host();
    ok = ok && O.to_host() == O2.to_host() && pv.size() == i.size() && !i.empty();
    for (size_t j = 0; ok && j < i.size(); ++j)
      ok = pv[j].index == (unsigned)i[j] && pv[j].value[0] == pts[i[j]][0] &&
           pv[j].value[1] == pts[i[j]][1] && pv[j].value[2] == pts[i[j]][2];
  }
  CHECK(ok, "attach_indices points: PairValueIndex rows = index rows with their values");

  std::vector<Box> bx(n);
  for (int j = 0; j < n; ++j)
    bx[j] = Box({{pts[j][0], pts[j][1], pts[j][2]}},
                {{pts[j][0] + 0.01f, pts[j][1] + 0.02f, pts[j][2] + 0.03f}});
  DeviceVector<Box> vb(bx);
  BVH btree(space, Experimental::attach_indices<int>(vb), fmt);
  DeviceVector<PairValueIndex<Box, int>> PB;
  btree.query(space, Experimental::make_intersects(q, 0.05f), PB, O2);
  BVH bplain(space, vb, fmt);
  bplain.query(space, Experimental::make_intersects(q, 0.05f), I, O);
  auto i = I.to_host();
  auto pb = PB.to_host();
  bool okb = O.to_host() == O2.to_host() && pb.size() == i.size() && !i.empty();
  for (size_t j = 0; okb && j < i.size(); ++j)
    okb = pb[j].index == i[j] && pb[j].value.minCorner()[0] == bx[i[j]].minCorner()[0] &&
          pb[j].value.maxCorner()[2] == bx[i[j]].maxCorner()[2];
  CHECK(okb, "attach_indices boxes: PairValueIndex<Box, int> rows");

  DeviceVector<Point> none;
  BVH empty_tree(space, Experimental::attach_indices(none), fmt);
  empty_tree.query(space, Experimental::make_intersects(q, 0.08f), P, O2);
  auto o = O2.to_host();
  CHECK(P.size() == 0 && o.size() == (size_t)nq + 1 && o.back() == 0,
        "attach_indices over no values: empty rows");

  bool threw = false;
  try
  {
    DeviceVector<PairValueIndex<Point>> wrong;
    btree.query(space, Experimental::make_intersects(q, 0.05f), wrong, O2);
  }
  catch (std::invalid_argument const &)
  {
    threw = true;
  }
  CHECK(threw, "PairValueIndex of another value type is refused");
}

// tstNeighborList.cpp:116-176: four collinear points, full list and half
// list expanded to full.
static void neighbor_lists(HipStream const &space)
{
  DeviceVector<Point> none;
  DeviceVector<int64_t> O;
  DeviceVector<int> I;
  Experimental::findFullNeighborList(space, none, INFINITY, O, I);
  CHECK(O.to_host() == std::vector<int64_t>{0} && I.size() == 0, "no point: offsets {0}");
  DeviceVector<Point> one(std::vector<Point>{{{0.f, 0.f, 0.f}}});
  Experimental::findHalfNeighborList(space, one, INFINITY, O, I);
  CHECK((O.to_host() == std::vector<int64_t>{0, 0}) && I.size() == 0, "one point: no pair");

  DeviceVector<Point> four(std::vector<Point>{
      {{0.f, 0.f, 0.f}}, {{1.f, 1.f, 1.f}}, {{2.f, 2.f, 2.f}}, {{3.f, 3.f, 3.f}}});
  struct Case
  {
    float r;
    std::vector<std::vector<int>> want;
  };
  std::vector<Case> cases = {
      {1.f, {{}, {}, {}, {}}},
      {2.f, {{1}, {0, 2}, {1, 3}, {2}}},
      {4.f, {{1, 2}, {0, 2, 3}, {0, 1, 3}, {1, 2}}},
      {6.f, {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}}},
  };
  for (auto &c : cases)
  {
    Experimental::findFullNeighborList(space, four, c.r, O, I);
    CHECK(rows_of(I, O) == rows(c.want), "full neighbour list");
    Experimental::findHalfNeighborList(space, four, c.r, O, I);
    Details::expandHalfToFull(space, O, I);
    CHECK(rows_of(I, O) == rows(c.want), "half list expanded to full");
  }
}

// test/tstQueryTreeRay.cpp: ten unit boxes on the diagonal.  Ordered rows
// are compared as sequences.
template <class Make>
static void rays(HipStream const &space, Make make)
{
  using Experimental::Ray;
  std::vector<Box> diag;
  for (int i = 0; i < 10; ++i)
    diag.push_back(Box({{(float)i, (float)i, (float)i}},
                       {{(float)i + 1, (float)i + 1, (float)i + 1}}));
  DeviceVector<Box> v(diag);
  auto index = make(space, v);
  DeviceVector<int> ind;
  DeviceVector<int64_t> off;
  // :60-87 intersects(ray)
  DeviceVector<Ray> r1(std::vector<Ray>{Ray({{0, 0, 0}}, {{.1f, .1f, .1f}})});
  index.query(space, intersects(r1), ind, off);
  CHECK(rows_of(ind, off) == rows({{0, 1, 2, 3, 4, 5, 6, 7, 8, 9}}), "rays: intersects");
  // :29-58 nearest(ray, 1)
  DeviceVector<Ray> r0(std::vector<Ray>{Ray({{0, 0, 0}}, {{.15f, .1f, 0.f}})});
  index.query(space, nearest(r0, 1), ind, off);
  CHECK(rows_of(ind, off) == rows({{0}}), "rays: nearest k = 1");
  DeviceVector<Ray> rb(std::vector<Ray>{Ray({{10, 10, 10}}, {{-.1f, -.1f, -.1f}})});
  index.query(space, nearest(rb, 3), ind, off);
  CHECK((ind.to_host() == std::vector<int>{9, 8, 7}), "rays: nearest k = 3 in order");
  // ordered ray queries: trees only, as in the reference
  if constexpr (std::is_same_v<typename Make::index_type, BoundingVolumeHierarchy>)
  {
    // :112-163 ordered intersections, forwards and backwards
    DeviceVector<Ray> r2(std::vector<Ray>{Ray({{0, 0, 0}}, {{.1f, .1f, .1f}}),
                                          Ray({{10, 10, 10}}, {{-.1f, -.1f, -.1f}})});
    index.query(space, Experimental::make_ordered_intersects(r2), ind, off);
    std::vector<int> want;
    for (int i = 0; i < 10; ++i)
      want.push_back(i);
    for (int i = 9; i >= 0; --i)
      want.push_back(i);
    CHECK(ind.to_host() == want && (off.to_host() == std::vector<int64_t>{0, 10, 20}),
          "rays: ordered intersections in order");
    index.query(space, nearest(r2, 3), ind, off);
    CHECK((ind.to_host() == std::vector<int>{0, 1, 2, 9, 8, 7}), "rays: nearest k = 3 in order");
    // :212-243 one leaf; :185-210 empty tree
    DeviceVector<Box> one(std::vector<Box>{Box({{0, 0, 0}}, {{1, 1, 1}})});
    auto leaf = make(space, one);
    DeviceVector<Ray> r3(std::vector<Ray>{Ray({{0, 0, 0}}, {{1, 1, 1}}),
                                          Ray({{4, 5, 6}}, {{7, 8, 9}})});
    leaf.query(space, Experimental::ordered_intersects(r3), ind, off);
    CHECK(rows_of(ind, off) == rows({{0}, {}}), "rays: one leaf");
    typename Make::index_type none;
    none.query(space, Experimental::ordered_intersects(r3), ind, off);
    CHECK(rows_of(ind, off) == rows({{}, {}}), "rays: empty tree");
    // the constructor normalises in double; zero components are +0
    Ray a({{0, 0, 0}}, {{3.f, -0.f, 4.f}});
    CHECK(a.direction()[0] == 0.6f && a.direction()[2] == 0.8f &&
              !std::signbit(a.direction()[1]),
          "rays: normalised direction");
  }
}

int main()
{
  HipStream space;
  for (NodeFormat fmt : {NodeFormat::FP32, NodeFormat::Q16})
  {
    degenerate(space, MakeBVH{fmt});
    callbacks_nearest(space, MakeBVH{fmt});
    manufactured(space, MakeBVH{fmt});
    brute_force(space, MakeBVH{fmt});
    rays(space, MakeBVH{fmt});
    attach_indices_values(space, fmt);
  }
  rays(space, MakeBruteForce{});
  policy(space, MakeBVH{NodeFormat::FP32});
  degenerate(space, MakeBruteForce{});
  callbacks_nearest(space, MakeBruteForce{});
  manufactured(space, MakeBruteForce{});
  brute_force(space, MakeBruteForce{});
  policy(space, MakeBruteForce{});
  neighbor_lists(space);
  std::printf("%d passed, %d failed\n", g_pass, g_fail);
  return g_fail == 0 ? 0 : 1;
}
