// test_heap_host.cpp -- the binary max-heap on (distance, index) pairs that
// the library keeps in a kNN output row for k > 64 (RowHeap in
// include/arborx_mi355x_device.hpp), run on the CPU:
//   * the insertion sequences of the reference's heap tests
//     (test/tstHeapOperations.cpp:25-36 and :52-101, pushHeap steps): the
//     heap array after each insertion equals the reference's expected array;
//   * top-k against std::partial_sort by (distance, index) on random streams
//     with repeated distances, with the heap property checked after every
//     insertion.
// Host code only (hipcc host pass); no GPU is used.
#include "arborx_mi355x_device.hpp"

#include <algorithm>
#include <cstdio>
#include <random>
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

static bool heap_ok(const std::vector<float> &d, const std::vector<int32_t> &i, int count)
{
  for (int j = 1; j < count; ++j)
  {
    const int par = (j - 1) / 2;
    if (axb::pair_less(d[par], i[par], d[j], i[j]))
      return false;
  }
  return true;
}

// pushes of `vals` (value j gets index j); after push s the first s + 1
// heap distances must equal want[s]
static void push_sequence(const std::vector<float> &vals,
                          const std::vector<std::vector<float>> &want, const char *what)
{
  const int k = (int)vals.size();
  std::vector<float> d(k, -1.f);
  std::vector<int32_t> idx(k, -1);
  axb::RowHeap h;
  h.init(k, idx.data(), d.data());
  bool ok = true;
  for (int s = 0; s < k; ++s)
  {
    h.insert(vals[s], s);
    ok = ok && heap_ok(d, idx, s + 1);
    if (!want[s].empty())
      ok = ok && std::equal(want[s].begin(), want[s].end(), d.begin());
  }
  CHECK(ok, what);
}

int main()
{
  // tstHeapOperations.cpp:25-36: {11, 5, 8, 3, 4} is a heap; pushing 15
  // gives {15, 5, 11, 3, 4, 8}
  push_sequence({11, 5, 8, 3, 4, 15},
                {{11}, {11, 5}, {11, 5, 8}, {11, 5, 8, 3}, {11, 5, 8, 3, 4}, {15, 5, 11, 3, 4, 8}},
                "push_heap example");
  // tstHeapOperations.cpp:52-101 (pushHeap steps of max_heap): 3 1 4 1 5 9 6
  push_sequence({3, 1, 4, 1, 5, 9, 6},
                {{3}, {3, 1}, {4, 1, 3}, {4, 1, 3, 1}, {5, 4, 3, 1, 1}, {9, 4, 5, 1, 1, 3},
                 {9, 4, 6, 1, 1, 3, 5}},
                "max_heap pushes");

  // top-k by (distance, index) against a plain partial sort
  std::mt19937 gen(7);
  bool ok_all = true, ok_heap = true;
  for (int t = 0; t < 400; ++t)
  {
    const int k = 1 + (int)(gen() % 150);
    const int n = (int)(gen() % 600);
    const int levels = 1 + (int)(gen() % 40); // few distinct distances: many ties
    std::vector<std::pair<float, int32_t>> all(n);
    for (int j = 0; j < n; ++j)
      all[j] = {(float)(gen() % levels) * 0.25f, (int32_t)(gen() % 100000)};
    const int m = std::min(k, n);
    std::vector<float> d(std::max(m, 1), -1.f);
    std::vector<int32_t> idx(std::max(m, 1), -1);
    axb::RowHeap h;
    h.init(m, idx.data(), d.data());
    for (auto &e : all)
    {
      if (h.accepts(e.first, e.second))
        h.insert(e.first, e.second);
      ok_heap = ok_heap && heap_ok(d, idx, std::min(h.count, m));
    }
    h.finish(m);
    std::partial_sort(all.begin(), all.begin() + m, all.end());
    for (int j = 0; j < m; ++j)
      ok_all = ok_all && d[j] == all[j].first && idx[j] == all[j].second;
  }
  CHECK(ok_heap, "heap property after every insertion");
  CHECK(ok_all, "top-k equals partial sort by (distance, index)");
  std::printf("%d passed, %d failed\n", g_pass, g_fail);
  return g_fail == 0 ? 0 : 1;
}
