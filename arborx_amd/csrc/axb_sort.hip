// axb_sort.hip -- stable LSD radix sort (onesweep) and device-wide scans.
//
// Replaces the reference's sortObjects -> sortByKey (misc/ArborX_SortUtils.hpp:
// 26-43, kokkos_ext/ArborX_KokkosExtSort.hpp:79-149: Kokkos BinSort on the
// CPU, rocThrust sort_by_key on HIP) and KokkosExt::exclusive_scan
// (kokkos_ext/ArborX_KokkosExtStdAlgorithms.hpp).
//
// Sort design for gfx950 (wave64):
//  * 8-bit digits; the bucket counts of every pass come from one histogram
//    computed by the key-generation kernel (no per-pass counting sweep);
//  * one kernel per pass: a block takes a 4096-key tile (tile id from an
//    atomic ticket so that every tile it waits on is already running), ranks
//    its keys per wavefront with 8 ballots per key (equal-digit lane masks),
//    publishes its per-digit counts, looks back over predecessor tiles for
//    the digit prefixes (decoupled look-back with bounded waits), stages the
//    tile in LDS in digit order and writes runs of equal digits contiguously;
//  * stability: a key's rank orders (wave, item, lane), which is its order in
//    the input, and tiles are prefixed in tile order.
#include "axb_internal.hpp"

namespace axb
{

namespace
{

constexpr uint32_t FLAG_AGG = 1u << 30;
constexpr uint32_t FLAG_PREFIX = 2u << 30;
constexpr uint32_t VALUE_MASK = (1u << 30) - 1u;

__global__ void k_fill_u32(uint32_t *p, uint32_t v, int64_t n)
{
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride)
    p[i] = v;
}
__global__ void k_iota_u32(uint32_t *p, int64_t n)
{
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride)
    p[i] = (uint32_t)i;
}
__global__ void k_fill_u64(uint64_t *p, uint64_t v, int64_t n)
{
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride)
    p[i] = v;
}

template <class K>
__global__ __launch_bounds__(256) void k_key_hist(const K *__restrict__ keys,
                                                  int64_t n, int passes,
                                                  uint32_t *__restrict__ hist)
{
  __shared__ uint32_t s_h[8 * 256];
  for (int j = threadIdx.x; j < passes * 256; j += blockDim.x)
    s_h[j] = 0;
  __syncthreads();
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride)
  {
    K k = keys[i];
    for (int p = 0; p < passes; ++p)
      atomicAdd(&s_h[p * 256 + (uint32_t)((k >> (8 * p)) & 255)], 1u);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < passes * 256; j += blockDim.x)
    if (s_h[j])
      atomicAdd(&hist[j], s_h[j]);
}

// Wave64 inclusive sum scan.
template <class T>
AXB_D T wave_inclusive_sum(T v)
{
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1)
  {
    T u = __shfl_up(v, o, 64);
    if (lane >= o)
      v += u;
  }
  return v;
}

// hist (passes x 256 counts) -> exclusive digit offsets in place, plus a
// per-pass flag telling whether every key has the same digit in that pass.
__global__ __launch_bounds__(256) void k_hist_scan(uint32_t *hist,
                                                   uint32_t *trivial,
                                                   int passes, int64_t n)
{
  __shared__ uint32_t s_w[4];
  const int d = threadIdx.x, lane = d & 63, wave = d >> 6;
  for (int p = 0; p < passes; ++p)
  {
    uint32_t c = hist[p * 256 + d];
    uint32_t inc = wave_inclusive_sum(c);
    if (lane == 63)
      s_w[wave] = inc;
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < wave; ++w)
      add += s_w[w];
    hist[p * 256 + d] = inc - c + add;
    if ((int64_t)c == n)
      trivial[p] = 1u;
    __syncthreads();
  }
}

template <class K>
__global__ __launch_bounds__(SORT_BLOCK) void k_onesweep(
    const K *__restrict__ kin, const uint32_t *__restrict__ vin,
    K *__restrict__ kout, uint32_t *__restrict__ vout, int64_t n, int shift,
    const uint32_t *__restrict__ digit_offset,
    const uint32_t *__restrict__ trivial, uint32_t *__restrict__ status,
    uint32_t *__restrict__ tile_counter, uint32_t *__restrict__ err)
{
  constexpr int WAVES = SORT_BLOCK / 64;
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_whist[WAVES][256];
  __shared__ uint32_t s_bstart[256];
  __shared__ uint32_t s_gbase[256];
  __shared__ uint32_t s_wsum[WAVES];
  __shared__ K s_keys[SORT_TILE];
  __shared__ uint32_t s_vals[SORT_TILE];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  if (*trivial)
  {
    // Every key has the same digit in this pass: the stable pass is the
    // identity; copy through.
    int64_t base = (int64_t)blockIdx.x * SORT_TILE;
    for (int j = tid; j < SORT_TILE; j += SORT_BLOCK)
    {
      int64_t i = base + j;
      if (i < n)
      {
        kout[i] = kin[i];
        vout[i] = vin ? vin[i] : (uint32_t)i;
      }
    }
    return;
  }

  if (tid == 0)
    s_tile = atomicAdd(tile_counter, 1u);
  for (int j = tid; j < WAVES * 256; j += SORT_BLOCK)
    (&s_whist[0][0])[j] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const int64_t tile_base = (int64_t)tile * SORT_TILE;
  const int tile_count = (int)((n - tile_base) < SORT_TILE ? (n - tile_base)
                                                           : SORT_TILE);

  // ---- load (wave w owns items [w*64*ITEMS, (w+1)*64*ITEMS) of the tile)
  K key[SORT_ITEMS];
  uint32_t val[SORT_ITEMS];
  uint32_t rank[SORT_ITEMS];
  uint32_t dig[SORT_ITEMS];
#pragma unroll
  for (int i = 0; i < SORT_ITEMS; ++i)
  {
    int local = wave * 64 * SORT_ITEMS + i * 64 + lane;
    int64_t gi = tile_base + local;
    if (local < tile_count)
    {
      key[i] = kin[gi];
      val[i] = vin ? vin[gi] : (uint32_t)gi;
      dig[i] = (uint32_t)((key[i] >> shift) & 255);
    }
    else
    {
      key[i] = 0;
      val[i] = 0;
      dig[i] = 256u;
    }
  }

  // ---- per-wave ranking by equal-digit lane masks
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int i = 0; i < SORT_ITEMS; ++i)
  {
    const uint32_t d = dig[i];
    const bool valid = d < 256u;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b)
    {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    const uint32_t before = (uint32_t)__popcll(m & lt_mask);
    uint32_t old = 0;
    if (valid)
      old = s_whist[wave][d];
    if (valid && before == 0)
      s_whist[wave][d] = old + (uint32_t)__popcll(m);
    rank[i] = old + before;
  }
  __syncthreads();

  // ---- per digit: wave prefixes and tile total; publish the aggregate
  const int d = tid;
  uint32_t total = 0;
#pragma unroll
  for (int w = 0; w < WAVES; ++w)
  {
    uint32_t c = s_whist[w][d];
    s_whist[w][d] = total;
    total += c;
  }
  uint32_t *my_status = status + (size_t)tile * 256 + d;
  if (tile == 0)
    st_agent(my_status, FLAG_PREFIX | total);
  else
    st_agent(my_status, FLAG_AGG | total);

  // ---- block-local digit starts (exclusive scan of totals over digits)
  uint32_t inc = wave_inclusive_sum(total);
  if (lane == 63)
    s_wsum[wave] = inc;
  __syncthreads();
  uint32_t add = 0;
  for (int w = 0; w < wave; ++w)
    add += s_wsum[w];
  const uint32_t bstart = inc - total + add;
  s_bstart[d] = bstart;
  __syncthreads();

  // ---- stage the tile in LDS in digit order
#pragma unroll
  for (int i = 0; i < SORT_ITEMS; ++i)
  {
    if (dig[i] < 256u)
    {
      uint32_t pos = s_bstart[dig[i]] + s_whist[wave][dig[i]] + rank[i];
      if (pos < (uint32_t)SORT_TILE)
      {
        s_keys[pos] = key[i];
        s_vals[pos] = val[i];
      }
    }
  }

  // ---- decoupled look-back for digit d
  uint32_t excl = 0;
  if (tile != 0)
  {
    int64_t j = (int64_t)tile - 1;
    uint32_t spins = 0;
    while (j >= 0)
    {
      uint32_t v = ld_agent(status + (size_t)j * 256 + d);
      uint32_t flag = v & ~VALUE_MASK;
      if (flag == 0u)
      {
        if (++spins >= SPIN_LIMIT)
        {
          raise_error(err, KERR_SPIN);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      excl += v & VALUE_MASK;
      if (flag == FLAG_PREFIX)
        break;
      --j;
    }
    st_agent(my_status, FLAG_PREFIX | (excl + total));
  }
  s_gbase[d] = digit_offset[d] + excl - bstart;
  __syncthreads();

  // ---- write out runs of equal digits
  for (int j = tid; j < tile_count; j += SORT_BLOCK)
  {
    K k = s_keys[j];
    uint32_t dd = (uint32_t)((k >> shift) & 255);
    uint32_t pos = s_gbase[dd] + (uint32_t)j;
    if ((int64_t)pos < n) // only a failed look-back could break this
    {
      kout[pos] = k;
      vout[pos] = s_vals[j];
    }
  }
}

// ------------------------------------------------------------ scans
// Reduce, then scan: (1) the sum of each tile of SCAN_TILE elements,
// (2) one block turns the tile sums into exclusive tile offsets, (3) each
// tile scans itself from its offset.  The input is read twice and no block
// waits on another.  A thread owns SCAN_ITEMS consecutive elements.

// Elements [base, base + SCAN_ITEMS) of a thread, 0 past n; running sums in
// v.  A whole, 16-byte aligned run is read with 16-byte loads.
template <class Tin>
AXB_D int64_t scan_load_items(const Tin *in, int64_t n, int64_t base, int64_t v[SCAN_ITEMS])
{
  static_assert(SCAN_ITEMS * sizeof(Tin) % 16 == 0, "runs of whole 16-byte words");
  Tin x[SCAN_ITEMS];
  if (base + SCAN_ITEMS <= n && ((uintptr_t)(in + base) & 15u) == 0)
  {
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i += 16 / (int)sizeof(Tin))
      *(uint4 *)&x[i] = *(const uint4 *)(in + base + i);
  }
  else
  {
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
      x[i] = (base + i < n) ? in[base + i] : Tin(0);
  }
  int64_t run = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i)
  {
    run += (int64_t)x[i];
    v[i] = run;
  }
  return run;
}

// Exclusive prefix of x over the block's threads, and the block's total
// (s_wsum: one slot per wavefront).
template <int BLOCK>
AXB_D int64_t block_exclusive_sum(int64_t x, int64_t *s_wsum, int64_t &total)
{
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t inc = wave_inclusive_sum(x);
  if (lane == 63)
    s_wsum[wave] = inc;
  __syncthreads();
  int64_t add = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < BLOCK / 64; ++w)
  {
    add += w < wave ? s_wsum[w] : 0;
    total += s_wsum[w];
  }
  return inc - x + add;
}

template <class Tin>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tile_sums(const Tin *in, int64_t n,
                                                               int64_t *__restrict__ sums)
{
  __shared__ int64_t s_wsum[SCAN_BLOCK / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  int64_t v[SCAN_ITEMS];
  const int64_t run = scan_load_items(in, n, base, v);
  int64_t total;
  block_exclusive_sum<SCAN_BLOCK>(run, s_wsum, total);
  if (threadIdx.x == 0)
    sums[blockIdx.x] = total;
}

// One block: sums[0, tiles) -> their exclusive prefix, in place.
constexpr int SCAN_OFFSETS_BLOCK = 1024;
__global__ __launch_bounds__(SCAN_OFFSETS_BLOCK) void k_scan_tile_offsets(int64_t *sums,
                                                                          int64_t tiles)
{
  __shared__ int64_t s_wsum[SCAN_OFFSETS_BLOCK / 64];
  int64_t carry = 0;
  for (int64_t c0 = 0; c0 < tiles; c0 += (int64_t)SCAN_OFFSETS_BLOCK * SCAN_ITEMS)
  {
    const int64_t base = c0 + (int64_t)threadIdx.x * SCAN_ITEMS;
    int64_t v[SCAN_ITEMS];
    const int64_t run = scan_load_items(sums, tiles, base, v);
    int64_t total;
    const int64_t pre = carry + block_exclusive_sum<SCAN_OFFSETS_BLOCK>(run, s_wsum, total);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
      if (base + i < tiles)
        sums[base + i] = pre + (i == 0 ? 0 : v[i - 1]);
    carry += total;
    __syncthreads(); // s_wsum is rewritten by the next chunk
  }
}

// flags (when given): bit i of byte e is set when input e * 8 + i exceeds
// limit (the CRS's rows that overflow their scratch row, read from the
// counts as they are scanned).
template <class Tin, class Tout, bool INCLUSIVE>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tiles(const Tin *in, Tout *out, int64_t n,
                                                           const int64_t *__restrict__ offs,
                                                           uint8_t *__restrict__ flags,
                                                           int64_t limit)
{
  static_assert(SCAN_ITEMS == 8, "one flag byte per thread");
  __shared__ int64_t s_wsum[SCAN_BLOCK / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  int64_t v[SCAN_ITEMS];
  const int64_t run = scan_load_items(in, n, base, v);
  if (flags && base < n)
  {
    uint32_t bits = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
      bits |= (uint32_t)(base + i < n && v[i] - (i == 0 ? 0 : v[i - 1]) > limit) << i;
    flags[base >> 3] = (uint8_t)bits;
  }
  int64_t total;
  const int64_t pre = offs[blockIdx.x] + block_exclusive_sum<SCAN_BLOCK>(run, s_wsum, total);
  Tout y[SCAN_ITEMS];
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i)
    y[i] = (Tout)(pre + (INCLUSIVE ? v[i] : (i == 0 ? 0 : v[i - 1])));
  if (base + SCAN_ITEMS <= n && ((uintptr_t)(out + base) & 15u) == 0)
  {
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i += 16 / (int)sizeof(Tout))
      *(uint4 *)(out + base + i) = *(const uint4 *)&y[i];
  }
  else
  {
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
      if (base + i < n)
        out[base + i] = y[i];
  }
}

inline int fill_grid(int64_t n)
{
  int64_t g = div_up(n, 256);
  return (int)(g < 4096 ? (g < 1 ? 1 : g) : 4096);
}

template <class K>
hipError_t radix_sort_pairs(K *keys0, K *keys1, uint32_t *vals0,
                            uint32_t *vals1, int64_t n, int passes,
                            uint32_t *hist, void *workspace, uint32_t *err,
                            hipStream_t stream)
{
  const int64_t tiles = div_up(n, SORT_TILE);
  // workspace layout: [hist passes*256][trivial passes][counters passes]
  //                   [status passes*tiles*256]
  uint32_t *ws = (uint32_t *)workspace;
  uint32_t *own_hist = ws;
  uint32_t *trivial = ws + passes * 256;
  uint32_t *counters = trivial + passes;
  uint32_t *status = counters + passes;
  const int64_t zero_words =
      (hist ? 0 : passes * 256) + 2 * passes + (int64_t)passes * tiles * 256;
  uint32_t *zero_from = hist ? trivial : own_hist;
  AXB_LAUNCH(k_fill_u32, dim3(fill_grid(zero_words)), dim3(256), 0,
                     stream, zero_from, 0u, zero_words);
  if (!hist)
  {
    hist = own_hist;
    int g = (int)(tiles < 2048 ? tiles : 2048);
    AXB_LAUNCH(k_key_hist<K>, dim3(g), dim3(256), 0, stream, keys0, n,
                       passes, hist);
  }
  AXB_LAUNCH(k_hist_scan, dim3(1), dim3(256), 0, stream, hist,
                     trivial, passes, n);
  K *kin = keys0, *kout = keys1;
  uint32_t *vin = nullptr, *vout = vals1;
  for (int p = 0; p < passes; ++p)
  {
    AXB_LAUNCH(k_onesweep<K>, dim3((unsigned)tiles), dim3(SORT_BLOCK),
                       0, stream, kin, vin, kout, vout, n, 8 * p,
                       hist + p * 256, trivial + p,
                       status + (size_t)p * tiles * 256, counters + p, err);
    K *kt = kin;
    kin = kout;
    kout = kt;
    vin = vout;
    vout = (vout == vals1) ? vals0 : vals1;
  }
  return hipGetLastError();
}

} // namespace

size_t sort_workspace_bytes(int64_t n, int passes, int)
{
  const int64_t tiles = div_up(n, SORT_TILE);
  return sizeof(uint32_t) *
         (size_t)(passes * 256 + 2 * passes + (int64_t)passes * tiles * 256);
}

hipError_t radix_sort_pairs_u64(uint64_t *keys0, uint64_t *keys1,
                                uint32_t *vals0, uint32_t *vals1, int64_t n,
                                int passes, uint32_t *hist, void *workspace,
                                uint32_t *err, hipStream_t stream)
{
  return radix_sort_pairs<uint64_t>(keys0, keys1, vals0, vals1, n, passes,
                                    hist, workspace, err, stream);
}
hipError_t radix_sort_pairs_u32(uint32_t *keys0, uint32_t *keys1,
                                uint32_t *vals0, uint32_t *vals1, int64_t n,
                                int passes, uint32_t *hist, void *workspace,
                                uint32_t *err, hipStream_t stream)
{
  return radix_sort_pairs<uint32_t>(keys0, keys1, vals0, vals1, n, passes,
                                    hist, workspace, err, stream);
}

size_t scan_workspace_bytes(int64_t n)
{
  return sizeof(int64_t) * (size_t)(div_up(n, SCAN_TILE) + 1);
}

// In place (out == in) is allowed: each element is read and written by the
// same thread.  n < 2^31 * SCAN_TILE (the grid's x limit).
template <class Tin, class Tout, bool INC>
static hipError_t scan_impl(const Tin *in, Tout *out, int64_t n, void *workspace, uint32_t *,
                            hipStream_t stream, uint8_t *flags = nullptr, int64_t limit = 0)
{
  if (n <= 0)
    return hipSuccess;
  const int64_t tiles = div_up(n, SCAN_TILE);
  int64_t *sums = (int64_t *)workspace;
  AXB_LAUNCH(k_scan_tile_sums<Tin>, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, stream, in, n,
             sums);
  AXB_LAUNCH(k_scan_tile_offsets, dim3(1), dim3(SCAN_OFFSETS_BLOCK), 0, stream, sums, tiles);
  AXB_LAUNCH((k_scan_tiles<Tin, Tout, INC>), dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, stream,
             in, out, n, sums, flags, limit);
  return hipGetLastError();
}

hipError_t exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, void *workspace,
                              uint32_t *err, hipStream_t stream, uint8_t *over_flags,
                              int64_t over_limit)
{
  return scan_impl<int64_t, int64_t, false>(in, out, n, workspace, err, stream, over_flags,
                                            over_limit);
}
hipError_t inclusive_scan_i32(const int32_t *in, int32_t *out, int64_t n,
                              void *workspace, uint32_t *err,
                              hipStream_t stream)
{
  return scan_impl<int32_t, int32_t, true>(in, out, n, workspace, err, stream);
}

hipError_t fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t stream)
{
  if (n <= 0)
    return hipSuccess;
  AXB_LAUNCH(k_fill_u32, dim3(fill_grid(n)), dim3(256), 0, stream, p,
                     v, n);
  return hipGetLastError();
}
hipError_t launch_iota_u32(uint32_t *p, int64_t n, hipStream_t stream)
{
  if (n <= 0)
    return hipSuccess;
  AXB_LAUNCH(k_iota_u32, dim3(fill_grid(n)), dim3(256), 0, stream, p, n);
  return hipGetLastError();
}
hipError_t fill_u64(uint64_t *p, uint64_t v, int64_t n, hipStream_t stream)
{
  if (n <= 0)
    return hipSuccess;
  AXB_LAUNCH(k_fill_u64, dim3(fill_grid(n)), dim3(256), 0, stream, p,
                     v, n);
  return hipGetLastError();
}

} // namespace axb
