// test_select_host.cpp -- the kNN row selection of
// include/arborx_mi355x_device.hpp (select_row, the code the collect kernel
// runs after its walk) run on the CPU against a plain sort.  A row holds c
// {index, d2 bits} entries; the answer is the m smallest by (distance,
// index) with distance = RN(sqrt(d2)), the nearest-query order of
// ArborX_LinearBVH.hpp / TreeTraversal.hpp (ties broken by index).  Rows are
// drawn so that some d2 differ by an ulp and round to the same distance, and
// some repeat exactly: the cases where ordering by d2 alone would be wrong.
// Host code only (hipcc host pass); no GPU is used.
#include "arborx_mi355x_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

static int g_fail = 0, g_pass = 0;

template <int K>
static void run(std::mt19937 &gen, int cases)
{
  std::uniform_real_distribution<float> U(0.f, 1.f);
  for (int t = 0; t < cases; ++t)
  {
    const int k = 1 + (int)(gen() % K);
    const int cap = 8 * k > 64 ? 8 * k : 64;
    const int m = k;
    const int c = m + (int)(gen() % (cap - m + 1));
    const int mode = (int)(gen() % 4); // 0 plain, 1 ulp neighbours, 2 repeats, 3 both
    std::vector<uint2> row(cap, make_uint2(0xdeadbeefu, 0x7f7fffffu));
    std::vector<uint32_t> ids(c);
    for (int j = 0; j < c; ++j)
      ids[j] = (uint32_t)(gen() % 1000000);
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    const int cc = (int)ids.size() < m ? (int)ids.size() : c;
    if (cc < m)
      continue;
    std::shuffle(ids.begin(), ids.end(), gen);
    std::vector<float> d2(cc);
    for (int j = 0; j < cc; ++j)
    {
      float v = 0.0001f + U(gen) * 0.01f;
      if ((mode & 1) && j > 0 && gen() % 3 == 0)
        v = std::nextafter(d2[gen() % j], 1.f); // often the same RN(sqrt)
      if ((mode & 2) && j > 0 && gen() % 4 == 0)
        v = d2[gen() % j]; // exact repeat
      d2[j] = v;
      uint32_t bits;
      std::memcpy(&bits, &v, 4);
      row[j] = make_uint2(ids[j], bits);
    }
    // expected: sort by (RN(sqrt(d2)), index)
    std::vector<std::pair<float, int32_t>> all(cc);
    for (int j = 0; j < cc; ++j)
      all[j] = {std::sqrt(d2[j]), (int32_t)ids[j]};
    std::sort(all.begin(), all.end());
    std::vector<int32_t> oi(m, -7);
    std::vector<float> od(m, -7.f);
    axb::select_row<K>(row.data(), 0u, cc, m, oi.data(), od.data(), 0);
    bool ok = true;
    for (int j = 0; j < m; ++j)
      ok = ok && oi[j] == all[j].second && std::memcmp(&od[j], &all[j].first, 4) == 0;
    // without distances
    std::vector<int32_t> oi2(m, -7);
    axb::select_row<K>(row.data(), 0u, cc, m, oi2.data(), nullptr, 0);
    ok = ok && oi2 == oi;
    if (ok)
      ++g_pass;
    else
    {
      ++g_fail;
      if (g_fail < 10)
        std::printf("FAIL K=%d k=%d c=%d mode=%d\n", K, k, cc, mode);
    }
  }
}

int main()
{
  std::mt19937 gen(12345u);
  run<1>(gen, 300);
  run<2>(gen, 300);
  run<4>(gen, 500);
  run<8>(gen, 1000);
  run<16>(gen, 3000);
  run<32>(gen, 1000);
  run<64>(gen, 300);
  std::printf("%d passed, %d failed\n", g_pass, g_fail);
  return g_fail == 0 ? 0 : 1;
}
