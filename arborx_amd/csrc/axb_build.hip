// axb_build.hip -- BVH construction kernels for gfx950.
//
// Replaces the reference's build (spatial/ArborX_LinearBVH.hpp:167-256):
//   scene box        TreeConstruction.hpp:26-39      -> k_scene_partial/final
//   Morton64 codes   SpaceFillingCurves.hpp:63-84    -> k_morton_hist (also
//                    builds the radix-sort digit histograms in the same pass)
//   sortObjects      SortUtils.hpp:26-43             -> axb_sort.hip
//   hierarchy        TreeConstruction.hpp:71-322     -> k_climb_local, k_climb_global
// and adds the pre-order linearisation the traversal kernels read
// (k_lin_*): the same topology stored depth-first so that a node's left child
// is the next record and its rope is `slot + skip`.
//
// Pre-order position of a node (no top-down pass needed): with c[j] = number
// of stored internal nodes whose first leaf is j and C its inclusive prefix
// sum, a stored internal node X with first leaf f, last leaf l and left-spine
// height h (stored nodes on the path from X down to leaf f, X included) is
// preceded by C[f] - h internal nodes and f leaves, and its subtree takes
// SL * (l - f + 1) + SI * (C[l] - C[f] + h) slots; leaf i is preceded by C[i]
// internal nodes and i leaves.  Nodes over at most FLAT_MAX leaves are not
// stored (h = 0); stored ones are the top part of every left spine.  c[] is
// written by the climb: the top of every left spine is the root or a right
// child, and its height is the spine length.
#include "axb_internal.hpp"

namespace axb
{

namespace
{

// ------------------------------------------------------------ scene box
template <int G>
__global__ __launch_bounds__(256) void k_scene_partial(const float *values,
                                                       int64_t n,
                                                       float *partials)
{
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
  float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride)
  {
    const float *v = values + i * (G == GEOM_POINTS ? 3 : 6);
    const float *w = v + (G == GEOM_POINTS ? 0 : 3);
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      lo[d] = kmin(lo[d], v[d]);
      hi[d] = kmax(hi[d], w[d]);
    }
  }
  // min/max are exact and order-free: any reduction tree gives the
  // reference's value.
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      lo[d] = kmin(lo[d], __shfl_xor(lo[d], o, 64));
      hi[d] = kmax(hi[d], __shfl_xor(hi[d], o, 64));
    }
  __shared__ float s[4][6];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    for (int d = 0; d < 3; ++d)
      s[wave][d] = lo[d], s[wave][3 + d] = hi[d];
  __syncthreads();
  if (threadIdx.x < 6)
  {
    const int d = threadIdx.x;
    float r = s[0][d];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
      r = d < 3 ? kmin(r, s[w][d]) : kmax(r, s[w][d]);
    partials[blockIdx.x * 6 + d] = r;
  }
}

// One block of 256 threads: strided partial reductions into LDS, then one
// thread per coordinate over the 256 values (min / max are exact in any
// order).
__global__ __launch_bounds__(256) void k_scene_final(const float *partials, int nblocks,
                                                     float *scene)
{
  __shared__ float s[6][256];
  const int t = threadIdx.x;
  float r[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int b = t; b < nblocks; b += 256)
    for (int d = 0; d < 6; ++d)
      r[d] = d < 3 ? kmin(r[d], partials[b * 6 + d]) : kmax(r[d], partials[b * 6 + d]);
  for (int d = 0; d < 6; ++d)
    s[d][t] = r[d];
  __syncthreads();
  if (t < 6)
  {
    float v = s[t][0];
    for (int j = 1; j < 256; ++j)
      v = t < 3 ? kmin(v, s[t][j]) : kmax(v, s[t][j]);
    scene[t] = v;
  }
}

// ------------------------------------------------- Morton64 + histograms
template <int G>
__global__ __launch_bounds__(256) void k_morton_hist(
    const float *__restrict__ values, int64_t n,
    const float *__restrict__ scene, uint64_t *__restrict__ keys,
    uint32_t *__restrict__ hist)
{
  __shared__ uint32_t s_h[8 * 256];
  for (int j = threadIdx.x; j < 8 * 256; j += blockDim.x)
    s_h[j] = 0;
  const float a0 = scene[0], a1 = scene[1], a2 = scene[2];
  const float b0 = scene[3], b1 = scene[4], b2 = scene[5];
  __syncthreads();
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride)
  {
    float cx, cy, cz;
    if (G == GEOM_POINTS)
    {
      cx = values[3 * i], cy = values[3 * i + 1], cz = values[3 * i + 2];
    }
    else
    {
      // Centroid.hpp:49-60
      const float *v = values + 6 * i;
      cx = (v[0] + v[3]) / 2, cy = (v[1] + v[4]) / 2, cz = (v[2] + v[5]) / 2;
    }
    uint64_t k = morton64(translate_and_scale(cx, a0, b0),
                          translate_and_scale(cy, a1, b1),
                          translate_and_scale(cz, a2, b2));
    keys[i] = k;
#pragma unroll
    for (int p = 0; p < 8; ++p)
      atomicAdd(&s_h[p * 256 + (uint32_t)((k >> (8 * p)) & 255)], 1u);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < 8 * 256; j += blockDim.x)
    if (s_h[j])
      atomicAdd(&hist[j], s_h[j]);
}

// ------------------------------------------------------------ climb
// TreeConstruction.hpp:127-170.
AXB_D int64_t apetrei_delta(const uint64_t *__restrict__ keys, int nint,
                            int i)
{
  if (i < 0 || i >= nint)
    return INT64_MAX;
  uint64_t x = keys[i] ^ keys[i + 1];
  return x ? (int64_t)(x - 1ull)
           : (int64_t)((uint64_t)INT64_MIN + (uint64_t)(i ^ (i + 1)) - 1ull);
}
// TreeConstruction.hpp:172-195.
AXB_D int32_t apetrei_rope(const uint64_t *__restrict__ keys, int nint, int n,
                           int range_right, int64_t delta_right)
{
  if (range_right != nint)
    return delta_right < apetrei_delta(keys, nint, range_right + 1)
               ? range_right + 1
               : range_right + 1 + n;
  return -1;
}

// Hand-off between the two threads that meet at a parent.  The data one
// thread leaves for the other (its node's record with box and spine height,
// or its leaf record) is written and read with 16-byte accesses that go
// through the XCD's L2 to memory (the sc1 policy that agent-scope atomics
// use), so they are coherent across the XCDs' L2s on their own; the writer
// waits for those stores to complete before its CAS.  This gives the
// release/acquire pairing of TreeConstruction.hpp:241-265 (CAS + load_fence)
// for exactly the words that cross threads, without the whole-cache
// writeback/invalidate that an acq_rel CAS at agent scope costs on gfx950.
// The record loads are issued together and waited for once.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
AXB_D uint4 to_uint4(u32x4 v) { return make_uint4(v.x, v.y, v.z, v.w); }
AXB_D void st_coh128(void *p, uint4 v)
{
  const u32x4 x = {v.x, v.y, v.z, v.w};
  asm volatile("global_store_dwordx4 %0, %1, off sc1" : : "v"(p), "v"(x) : "memory");
}
AXB_D uint4 ld_coh128(const void *p)
{
  u32x4 a;
  asm volatile("global_load_dwordx4 %0, %1, off sc1\n\t"
               "s_waitcnt vmcnt(0)"
               : "=&v"(a)
               : "v"(p)
               : "memory");
  return to_uint4(a);
}
AXB_D void ld_coh128x2(const void *p, uint4 &a, uint4 &b)
{
  u32x4 x, y;
  asm volatile("global_load_dwordx4 %0, %2, off sc1\n\t"
               "global_load_dwordx4 %1, %2, off offset:16 sc1\n\t"
               "s_waitcnt vmcnt(0)"
               : "=&v"(x), "=&v"(y)
               : "v"(p)
               : "memory");
  a = to_uint4(x), b = to_uint4(y);
}
AXB_D void ld_coh128x3(const void *p, uint4 &a, uint4 &b, uint4 &c)
{
  u32x4 x, y, z;
  asm volatile("global_load_dwordx4 %0, %3, off sc1\n\t"
               "global_load_dwordx4 %1, %3, off offset:16 sc1\n\t"
               "global_load_dwordx4 %2, %3, off offset:32 sc1\n\t"
               "s_waitcnt vmcnt(0)"
               : "=&v"(x), "=&v"(y), "=&v"(z)
               : "v"(p)
               : "memory");
  a = to_uint4(x), b = to_uint4(y), c = to_uint4(z);
}

AXB_D int cas_handoff(int32_t *p, int32_t expected, int32_t desired)
{
  // every coherent store issued so far has completed before the flag moves
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  return expected;
}

// Reads the box (and spine height) of a completed sibling.
template <int G>
AXB_D void read_sibling(bool is_leaf, int j, const float4 *leafrec, const INode *inodes,
                        Box6 &b, int &h)
{
  if (is_leaf)
  {
    if (G == GEOM_POINTS)
    {
      const uint4 a = ld_coh128(leafrec + j);
      b.lo[0] = b.hi[0] = bits_f(a.x), b.lo[1] = b.hi[1] = bits_f(a.y);
      b.lo[2] = b.hi[2] = bits_f(a.z);
    }
    else
    {
      uint4 a, c;
      ld_coh128x2(leafrec + 2 * j, a, c);
      b.lo[0] = bits_f(a.x), b.lo[1] = bits_f(a.y), b.lo[2] = bits_f(a.z);
      b.hi[0] = bits_f(c.x), b.hi[1] = bits_f(c.y), b.hi[2] = bits_f(c.z);
    }
    h = 0;
  }
  else
  {
    // INode: {lo xyz, left_child} {hi xyz, rope} {first, last, h, pad}
    uint4 a, c, e;
    ld_coh128x3(inodes + j, a, c, e);
    b.lo[0] = bits_f(a.x), b.lo[1] = bits_f(a.y), b.lo[2] = bits_f(a.z);
    b.hi[0] = bits_f(c.x), b.hi[1] = bits_f(c.y), b.hi[2] = bits_f(c.z);
    h = (int)e.z;
  }
}

// The climb of TreeConstruction.hpp:197-311 (one thread per leaf climbs
// toward the root; the first thread to reach a parent stops, the second
// finishes it) in two kernels.  k_climb_local: a block of CLIMB_LEAVES threads
// takes as many consecutive leaves and climbs as long as the parent's split
// lies between two of its own leaves, with the hand-off flags, the deltas of
// its splits and the boxes of its nodes in LDS; a node whose parent split lies
// between two blocks (at most one per block side) is handed to
// k_climb_global, which continues with the flags in device memory.  Both
// write the INode records, the sorted leaf records, the spine data for the
// pre-order positions and (optionally) the reference-form ropes.
constexpr int CLIMB_LEAVES = 256;

template <int G>
AXB_D void load_leaf(const float *__restrict__ values, uint32_t orig, float4 *leafrec,
                     int i0, Box6 &bv)
{
  if (G == GEOM_POINTS)
  {
    const float *v = values + 3 * (int64_t)orig;
    float x = v[0], y = v[1], z = v[2];
    // Box{} expanded by the point (Expand.hpp:42-57)
    bv.lo[0] = kmin(FLT_MAX, x), bv.lo[1] = kmin(FLT_MAX, y), bv.lo[2] = kmin(FLT_MAX, z);
    bv.hi[0] = kmax(-FLT_MAX, x), bv.hi[1] = kmax(-FLT_MAX, y), bv.hi[2] = kmax(-FLT_MAX, z);
    ((uint4 *)leafrec)[i0] = make_uint4(f_bits(x), f_bits(y), f_bits(z), ~orig);
  }
  else
  {
    const float *v = values + 6 * (int64_t)orig;
    for (int d = 0; d < 3; ++d)
      bv.lo[d] = kmin(FLT_MAX, v[d]), bv.hi[d] = kmax(-FLT_MAX, v[3 + d]);
    ((uint4 *)leafrec)[2 * i0] =
        make_uint4(f_bits(bv.lo[0]), f_bits(bv.lo[1]), f_bits(bv.lo[2]), ~orig);
    ((uint4 *)leafrec)[2 * i0 + 1] =
        make_uint4(f_bits(bv.hi[0]), f_bits(bv.hi[1]), f_bits(bv.hi[2]), 0u);
  }
}

// Rope of a node (TreeConstruction.hpp:172-195) from the deltas of its last
// split and the next one.
AXB_D int32_t rope_of(int nint, int n, int range_right, int64_t delta_right,
                      int64_t delta_next)
{
  if (range_right == nint)
    return -1;
  return delta_right < delta_next ? range_right + 1 : range_right + 1 + n;
}

template <int G, bool ROPES>
__global__ __launch_bounds__(CLIMB_LEAVES) void k_climb_local(
    const float *__restrict__ values, int n, const uint64_t *__restrict__ keys,
    const uint32_t *__restrict__ perm, int32_t *__restrict__ ranges,
    INode *__restrict__ inodes, float4 *__restrict__ leafrec,
    int32_t *__restrict__ leaf_rope, int32_t *__restrict__ c, int32_t *__restrict__ top,
    int4 *__restrict__ pending, int32_t *__restrict__ spine_count,
    int32_t *__restrict__ spine_list)
{
  constexpr int B = CLIMB_LEAVES;
  constexpr int LW = G == GEOM_POINTS ? 3 : 6; // words of a leaf box in LDS
  __shared__ int64_t s_delta[B + 2];           // delta(base - 1 + j)
  __shared__ int32_t s_range[B];               // flag of split base + j
  __shared__ float s_node[7][B];               // lo, hi, spine height of node base + j
  __shared__ float s_leaf[LW][B];              // box of leaf base + j
  __shared__ int4 s_pend[2];                   // nodes at the left / right block split
  __shared__ int32_t s_sp[B];                  // first leaves of stored left spines
  __shared__ int32_t s_nsp;
  const int t = threadIdx.x;
  const int base = blockIdx.x * B;
  const int m = min(B, n - base);
  const int nint = n - 1;
  for (int j = t; j < B + 2; j += B)
    s_delta[j] = apetrei_delta(keys, nint, base - 1 + j);
  s_range[t] = -1;
  if (t < 2)
    s_pend[t] = make_int4(-1, 0, 0, 0);
  if (t == 0)
    s_nsp = 0;
  Box6 bv;
  if (t < m)
  {
    load_leaf<G>(values, perm[base + t], leafrec, base + t, bv);
    for (int d = 0; d < 3; ++d)
    {
      s_leaf[d][t] = bv.lo[d];
      if (G == GEOM_BOXES)
        s_leaf[3 + d][t] = bv.hi[d];
    }
  }
  __syncthreads();
  if (t < m)
  {
    const int i0 = base + t;
    int range_left = i0, range_right = i0;
    int64_t dl = s_delta[t], dr = s_delta[t + 1];
    if (ROPES)
      leaf_rope[i0] = rope_of(nint, n, i0, dr, s_delta[t + 2]);
    int cur = i0;  // reference encoding of the current node
    int cur_h = 0; // left-spine height of the current node (0 for a leaf)
    do
    {
      const bool is_left = dr < dl;
      const int ap = is_left ? range_right : range_left - 1;
      if (ap < base || ap > base + m - 2)
      {
        // the parent's split lies between two blocks
        s_pend[is_left ? 1 : 0] = make_int4(cur, cur_h, range_left, range_right);
        break;
      }
      // this node's LDS record is written before the flag moves
      __builtin_amdgcn_s_waitcnt(0xc07f); // lgkmcnt(0)
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      int32_t old = -1;
      __hip_atomic_compare_exchange_strong(&s_range[ap - base], &old,
                                           is_left ? range_left : range_right,
                                           __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      if (old == -1)
        break;
      int left_child, h_left, right_node, h_right, sib;
      bool sib_leaf;
      if (is_left)
      {
        range_right = old;
        left_child = cur;
        h_left = cur_h;
        sib = ap + 1;
        sib_leaf = sib == range_right;
        dr = s_delta[range_right - base + 1];
        right_node = sib_leaf ? sib : sib + n;
      }
      else
      {
        range_left = old;
        sib = ap;
        sib_leaf = sib == range_left;
        dl = s_delta[range_left - base];
        left_child = sib_leaf ? sib : sib + n;
        right_node = cur;
        h_right = cur_h;
      }
      const int js = sib - base;
      int sh = 0;
      for (int d = 0; d < 3; ++d)
      {
        const float slo = sib_leaf ? s_leaf[d][js] : s_node[d][js];
        const float shi =
            sib_leaf ? s_leaf[G == GEOM_POINTS ? d : 3 + d][js] : s_node[3 + d][js];
        bv.lo[d] = kmin(bv.lo[d], slo);
        bv.hi[d] = kmax(bv.hi[d], shi);
      }
      if (!sib_leaf)
        sh = __float_as_int(s_node[6][js]);
      if (is_left)
        h_right = sh;
      else
        h_left = sh;
      const int kp = dr < dl ? range_right : range_left;
      // spine height among stored nodes; 0 = not stored (at most FLAT_MAX
      // leaves); stored nodes are the top part of every left spine
      const int hp =
          range_right - range_left + 1 > FLAT_MAX ? (left_child < n ? 0 : h_left) + 1 : 0;
      const int32_t rope =
          ROPES ? rope_of(nint, n, range_right, dr, s_delta[range_right - base + 2]) : -1;
      uint4 *w = (uint4 *)(inodes + kp);
      w[0] = make_uint4(f_bits(bv.lo[0]), f_bits(bv.lo[1]), f_bits(bv.lo[2]),
                        (uint32_t)left_child);
      w[1] = make_uint4(f_bits(bv.hi[0]), f_bits(bv.hi[1]), f_bits(bv.hi[2]), (uint32_t)rope);
      w[2] = make_uint4((uint32_t)range_left, (uint32_t)range_right, (uint32_t)hp, 0u);
      const int jk = kp - base;
      for (int d = 0; d < 3; ++d)
        s_node[d][jk] = bv.lo[d], s_node[3 + d][jk] = bv.hi[d];
      s_node[6][jk] = __int_as_float(hp);
      if (right_node >= n)
      {
        // the right child tops a left spine starting at leaf ap + 1
        c[ap + 1] = h_right;
        top[ap + 1] = right_node - n;
        if (h_right > 0)
          s_sp[atomicAdd(&s_nsp, 1)] = ap + 1;
      }
      cur = kp + n;
      cur_h = hp;
      if (kp == 0)
      {
        c[0] = hp;
        top[0] = 0;
        if (hp > 0)
          s_sp[atomicAdd(&s_nsp, 1)] = 0;
      }
    } while (cur != n);
  }
  __syncthreads();
  const int nsp = s_nsp;
  if (t < nsp)
    spine_list[base + t] = s_sp[t];
  if (t == 0)
    spine_count[blockIdx.x] = nsp;
  // first arrivals at the block's own splits (the global climb completes
  // their parents); the split after the block starts unclaimed
  if (t < m - 1)
    ranges[base + t] = s_range[t];
  else if (t == m - 1 && base + t < nint)
    ranges[base + t] = -1;
  if (t < 2)
    pending[2 * blockIdx.x + t] = s_pend[t];
}

// Continues the climb of the nodes k_climb_local handed over (one thread per
// hand-off slot), with the flags in device memory.
template <int G, bool ROPES>
__global__ __launch_bounds__(256) void k_climb_global(
    const int4 *__restrict__ pending, int npending, int n,
    const uint64_t *__restrict__ keys, int32_t *ranges, INode *inodes, float4 *leafrec,
    int32_t *__restrict__ c, int32_t *__restrict__ top, int32_t *__restrict__ spine_count,
    int32_t *__restrict__ spine_list)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npending)
    return;
  const int4 e = pending[j];
  if (e.x < 0)
    return;
  const int nint = n - 1;
  int cur = e.x, cur_h = e.y, range_left = e.z, range_right = e.w;
  Box6 bv;
  int unused_h;
  read_sibling<G>(cur < n, cur < n ? cur : cur - n, leafrec, inodes, bv, unused_h);
  int64_t dl = apetrei_delta(keys, nint, range_left - 1);
  int64_t dr = apetrei_delta(keys, nint, range_right);
  do
  {
    const bool is_left = dr < dl;
    int ap, left_child, h_left, right_node, h_right;
    Box6 sb;
    if (is_left)
    {
      ap = range_right;
      int old = cas_handoff(&ranges[ap], -1, range_left);
      if (old == -1)
        break;
      range_right = old;
      left_child = cur;
      h_left = cur_h;
      const int rc = ap + 1;
      const bool rc_leaf = (rc == range_right);
      dr = apetrei_delta(keys, nint, range_right);
      read_sibling<G>(rc_leaf, rc, leafrec, inodes, sb, h_right);
      right_node = rc_leaf ? rc : rc + n;
    }
    else
    {
      ap = range_left - 1;
      int old = cas_handoff(&ranges[ap], -1, range_right);
      if (old == -1)
        break;
      range_left = old;
      const int lc = ap;
      const bool lc_leaf = (lc == range_left);
      dl = apetrei_delta(keys, nint, range_left - 1);
      read_sibling<G>(lc_leaf, lc, leafrec, inodes, sb, h_left);
      left_child = lc_leaf ? lc : lc + n;
      right_node = cur;
      h_right = cur_h;
    }
    for (int d = 0; d < 3; ++d)
    {
      bv.lo[d] = kmin(bv.lo[d], sb.lo[d]);
      bv.hi[d] = kmax(bv.hi[d], sb.hi[d]);
    }
    const int kp = dr < dl ? range_right : range_left;
    const int hp =
        range_right - range_left + 1 > FLAT_MAX ? (left_child < n ? 0 : h_left) + 1 : 0;
    const int32_t rope = ROPES ? apetrei_rope(keys, nint, n, range_right, dr) : -1;
    uint4 *w = (uint4 *)(inodes + kp);
    st_coh128(w, make_uint4(f_bits(bv.lo[0]), f_bits(bv.lo[1]), f_bits(bv.lo[2]),
                            (uint32_t)left_child));
    st_coh128(w + 1, make_uint4(f_bits(bv.hi[0]), f_bits(bv.hi[1]), f_bits(bv.hi[2]),
                                (uint32_t)rope));
    st_coh128(w + 2, make_uint4((uint32_t)range_left, (uint32_t)range_right, (uint32_t)hp, 0u));
    // first leaves of stored left spines topped here (the right child's and,
    // at the root, the root's), appended to the list of the local-climb
    // block holding the leaf
    if (right_node >= n)
    {
      c[ap + 1] = h_right;
      top[ap + 1] = right_node - n;
      if (h_right > 0)
      {
        const int cb = (ap + 1) / CLIMB_LEAVES;
        spine_list[cb * CLIMB_LEAVES + atomicAdd(&spine_count[cb], 1)] = ap + 1;
      }
    }
    cur = kp + n;
    cur_h = hp;
    if (kp == 0)
    {
      c[0] = hp;
      top[0] = 0;
      if (hp > 0)
        spine_list[atomicAdd(&spine_count[0], 1)] = 0;
    }
  } while (cur != n);
}

// ------------------------------------------------------------ linearise
template <int G>
__global__ __launch_bounds__(256) void k_lin_fp32(const INode *__restrict__ inodes,
                                                  int nint,
                                                  const int32_t *__restrict__ C,
                                                  uint4 *__restrict__ slots, uint32_t total,
                                                  uint32_t *err)
{
  constexpr uint32_t SL = leaf_slots<G>(), SI = 2;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nint)
    return;
  const INode r = inodes[k];
  if (r.h == 0) // not stored: its leaves follow each other
    return;
  const uint32_t pos = SI * (uint32_t)(C[r.first] - r.h) + SL * (uint32_t)r.first;
  const uint32_t len = (uint32_t)(r.last - r.first + 1);
  const uint32_t skip = SL * len + SI * (uint32_t)(C[r.last] - C[r.first] + r.h);
  if (pos + SI > total || pos + skip > total)
  {
    raise_error(err, KERR_LAYOUT);
    return;
  }
  slots[pos] = make_uint4(f_bits(r.lo[0]), f_bits(r.lo[1]), f_bits(r.lo[2]), skip);
  slots[pos + 1] = make_uint4(f_bits(r.hi[0]), f_bits(r.hi[1]), f_bits(r.hi[2]), 0u);
}

// Q16: one thread per stored left spine (rooted at the root or at a right
// child, whose frame is the scene box), a wavefront per list of spines that
// the climb made for CLIMB_LEAVES consecutive leaves; walking down the spine,
// each left child is encoded relative to the decoded box of its parent.
template <int G>
__global__ __launch_bounds__(256) void k_lin_q16(const INode *__restrict__ inodes,
                                                 int n,
                                                 const int32_t *__restrict__ spine_count,
                                                 const int32_t *__restrict__ spine_list,
                                                 int nlists,
                                                 const int32_t *__restrict__ top,
                                                 const int32_t *__restrict__ C,
                                                 const float *__restrict__ scene,
                                                 uint4 *__restrict__ slots,
                                                 uint32_t total, uint32_t *err)
{
  constexpr uint32_t SL = leaf_slots<G>(), SI = 1;
  const int cb = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (cb >= nlists)
    return;
  const int cnt = spine_count[cb];
  if (cnt > CLIMB_LEAVES)
  {
    raise_error(err, KERR_LAYOUT);
    return;
  }
  for (int i = threadIdx.x & 63; i < cnt; i += 64)
  {
    const int j = spine_list[cb * CLIMB_LEAVES + i];
    if (j < 0 || j >= n)
    {
      raise_error(err, KERR_LAYOUT);
      return;
    }
    Box6 f;
    for (int d = 0; d < 3; ++d)
      f.lo[d] = scene[d], f.hi[d] = scene[3 + d];
    const uint32_t cj = (uint32_t)C[j]; // every node of the spine has first leaf j
    INode r = inodes[top[j]];
    while (r.h != 0) // a node with h = 0 and the rest of the spine are not stored
    {
      // the left child's record is requested before this node is encoded
      const bool more = r.left_child >= n;
      INode rn;
      if (more)
        rn = inodes[r.left_child - n];
      uint32_t q[6];
      Box6 dec;
      for (int d = 0; d < 3; ++d)
      {
        if (r.lo[d] < f.lo[d] || r.hi[d] > f.hi[d])
          raise_error(err, KERR_FRAME);
        q16_encode_axis(f.lo[d], f.hi[d], r.lo[d], r.hi[d], q[d], q[3 + d]);
        const float s = q16_scale(f.lo[d], f.hi[d]);
        dec.lo[d] = q16_lo(f.lo[d], s, q[d]);
        dec.hi[d] = q16_lo(f.lo[d], s, q[3 + d]);
      }
      const uint32_t pos = SI * (cj - (uint32_t)r.h) + SL * (uint32_t)j;
      const uint32_t len = (uint32_t)(r.last - j + 1);
      const uint32_t skip = SL * len + SI * ((uint32_t)C[r.last] - cj + (uint32_t)r.h);
      if (r.first != j || pos + SI > total || pos + skip > total)
      {
        raise_error(err, KERR_LAYOUT);
        return;
      }
      slots[pos] = make_uint4(q[0] | (q[1] << 16), q[2] | (q[3] << 16),
                              q[4] | (q[5] << 16), skip);
      if (!more)
        break;
      r = rn;
      f = dec;
    }
  }
}

template <int G, int SI>
__global__ __launch_bounds__(256) void k_lin_leaves(const float4 *__restrict__ leafrec,
                                                    int n,
                                                    const int32_t *__restrict__ C,
                                                    uint4 *__restrict__ slots, uint32_t total,
                                                    uint32_t *err)
{
  constexpr uint32_t SL = leaf_slots<G>();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const uint32_t pos = SL * (uint32_t)i + (uint32_t)SI * (uint32_t)C[i];
  if (pos + SL > total)
  {
    raise_error(err, KERR_LAYOUT);
    return;
  }
  const uint4 *src = (const uint4 *)leafrec;
  slots[pos] = src[SL * i];
  if (SL == 2)
    slots[pos + 1] = src[SL * i + 1];
}

template <int G>
__global__ void k_single_leaf(const float *values, uint4 *slots, float *scene)
{
  if (threadIdx.x != 0)
    return;
  const float *v = values;
  const float *w = values + (G == GEOM_POINTS ? 0 : 3);
  float lo[3], hi[3];
  for (int d = 0; d < 3; ++d)
    lo[d] = kmin(FLT_MAX, v[d]), hi[d] = kmax(-FLT_MAX, w[d]);
  slots[0] = make_uint4(f_bits(lo[0]), f_bits(lo[1]), f_bits(lo[2]), ~0u);
  if (G == GEOM_BOXES)
    slots[1] = make_uint4(f_bits(hi[0]), f_bits(hi[1]), f_bits(hi[2]), 0u);
  for (int d = 0; d < 3; ++d)
    scene[d] = lo[d], scene[3 + d] = hi[d];
}

// Zeroes the SLOT_PAD slots after the last stored record (read as slot
// p + 1 by the box walks after a record at the end of the array).
template <int SL, int SI>
__global__ void k_pad_tail(uint4 *slots, const int32_t *C, int n)
{
  if (threadIdx.x < SLOT_PAD)
  {
    const uint32_t total = SL * (uint32_t)n + SI * (uint32_t)C[n - 1];
    slots[total + threadIdx.x] = make_uint4(0u, 0u, 0u, 0u);
  }
}

__global__ void k_export_topology(const INode *inodes, int nint,
                                  int32_t *left_child, int32_t *rope,
                                  float *boxes)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nint)
    return;
  const INode r = inodes[k];
  if (left_child)
    left_child[k] = r.left_child;
  if (rope)
    rope[k] = r.rope;
  if (boxes)
    for (int d = 0; d < 3; ++d)
      boxes[6 * k + d] = r.lo[d], boxes[6 * k + 3 + d] = r.hi[d];
}

__global__ void k_build_status(const float *scene, const int32_t *C, int64_t n,
                               uint32_t *status)
{
  const int t = threadIdx.x;
  if (t == 0)
    status[1] = n > 1 ? (uint32_t)C[n - 1] : 0u;
  else if (t < 7)
    status[1 + t] = f_bits(scene[t - 1]);
}

inline unsigned grid_for(int64_t n, int block = 256)
{
  return (unsigned)div_up(n, block);
}

} // namespace

// ------------------------------------------------------------ host driver

// Upper bound of the node array (every internal node stored); the actual
// size is SL * n + SI * C[n - 1] (BuildOut::total_slots after the build).
int64_t total_slots_for(int64_t n, int geometry, int node_format)
{
  if (n == 0)
    return 0;
  const int sl = geometry == GEOM_POINTS ? 1 : 2;
  const int si = node_format == NODES_FP32 ? 2 : 1;
  return sl * n + si * (n - 1);
}

// Builds the tree; all device buffers in `out` are allocated here with
// dalloc on `stream`.  keep_topology keeps keys/perm/leaf_rope/inodes.
hipError_t scene_box(const float *values, int64_t n, int geometry, float *scene,
                     hipStream_t stream)
{
  const int partial_blocks = (int)std::min<int64_t>(1024, div_up(n > 0 ? n : 1, 256));
  float *partials = (float *)dalloc(partial_blocks * 6 * sizeof(float), stream);
  if (!partials)
    return hipErrorOutOfMemory;
  if (geometry == GEOM_POINTS)
    AXB_LAUNCH(k_scene_partial<GEOM_POINTS>, dim3(partial_blocks), dim3(256), 0, stream, values,
               n, partials);
  else
    AXB_LAUNCH(k_scene_partial<GEOM_BOXES>, dim3(partial_blocks), dim3(256), 0, stream, values,
               n, partials);
  AXB_LAUNCH(k_scene_final, dim3(1), dim3(256), 0, stream, partials, partial_blocks, scene);
  const hipError_t e = hipGetLastError();
  dfree(partials, stream);
  return e;
}

hipError_t build_tree(const float *values, int64_t n, int geometry,
                      int node_format, bool keep_topology, uint32_t *err,
                      hipStream_t stream, BuildOut &out)
{
  out = BuildOut{};
  // allocated for every internal node stored; the caller sets total_slots to
  // SL * n + SI * C[n - 1] once C is on the host
  out.max_slots = total_slots_for(n, geometry, node_format);
  out.total_slots = out.max_slots;
  out.scene = (float *)dalloc(6 * sizeof(float), stream);
  out.slots = (uint4 *)dalloc((size_t)(out.max_slots + SLOT_PAD) * sizeof(uint4), stream);
  if (!out.scene || !out.slots)
    return hipErrorOutOfMemory;
  // zero slots after the last record (kernels read ahead up to SLOT_PAD
  // slots and check the record kind)
  hipError_t e0 = fill_u32((uint32_t *)(out.slots + out.max_slots), 0u, 4 * SLOT_PAD, stream);
  if (e0 != hipSuccess)
    return e0;
  if (n == 0)
  {
    const float empty[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
    return hipMemcpyAsync(out.scene, empty, sizeof(empty), hipMemcpyHostToDevice, stream);
  }
  if (n == 1)
  {
    out.C = (int32_t *)dalloc(sizeof(int32_t), stream);
    e0 = fill_u32((uint32_t *)out.C, 0u, 1, stream);
    if (e0 != hipSuccess)
      return e0;
    if (geometry == GEOM_POINTS)
      AXB_LAUNCH(k_single_leaf<GEOM_POINTS>, dim3(1), dim3(64), 0, stream, values,
                         out.slots, out.scene);
    else
      AXB_LAUNCH(k_single_leaf<GEOM_BOXES>, dim3(1), dim3(64), 0, stream, values,
                         out.slots, out.scene);
    return hipGetLastError();
  }
  const int ni = (int)n, nint = ni - 1;
  const int leafrec_f4 = geometry == GEOM_POINTS ? 1 : 2;

  // ---- workspace
  const int partial_blocks = (int)std::min<int64_t>(1024, div_up(n, 256));
  size_t sort_ws = sort_workspace_bytes(n, 8, 8);
  size_t scan_ws = scan_workspace_bytes(n);
  uint64_t *keys0 = (uint64_t *)dalloc(n * sizeof(uint64_t), stream);
  uint64_t *keys1 = (uint64_t *)dalloc(n * sizeof(uint64_t), stream);
  uint32_t *vals0 = (uint32_t *)dalloc(n * sizeof(uint32_t), stream);
  uint32_t *vals1 = (uint32_t *)dalloc(n * sizeof(uint32_t), stream);
  uint32_t *hist = (uint32_t *)dalloc(8 * 256 * sizeof(uint32_t), stream);
  float *partials = (float *)dalloc(partial_blocks * 6 * sizeof(float), stream);
  void *ws = dalloc(std::max(sort_ws, scan_ws), stream);
  int32_t *ranges = (int32_t *)dalloc(nint * sizeof(int32_t), stream);
  INode *inodes = (INode *)dalloc(nint * sizeof(INode), stream);
  float4 *leafrec = (float4 *)dalloc((size_t)n * leafrec_f4 * sizeof(float4), stream);
  int32_t *c = (int32_t *)dalloc(n * sizeof(int32_t), stream);
  int32_t *top = (int32_t *)dalloc(n * sizeof(int32_t), stream);
  const int64_t nlists = div_up(n, CLIMB_LEAVES);
  int4 *pending = (int4 *)dalloc(2 * nlists * sizeof(int4), stream);
  int32_t *spine_count = (int32_t *)dalloc(nlists * sizeof(int32_t), stream);
  int32_t *spine_list = (int32_t *)dalloc(nlists * CLIMB_LEAVES * sizeof(int32_t), stream);
  out.C = (int32_t *)dalloc(n * sizeof(int32_t), stream);
  int32_t *leaf_rope = keep_topology ? (int32_t *)dalloc(n * sizeof(int32_t), stream) : nullptr;
  if (!keys0 || !keys1 || !vals0 || !vals1 || !hist || !partials || !ws || !ranges ||
      !inodes || !leafrec || !c || !top || !pending || !spine_count ||
      !spine_list || !out.C || (keep_topology && !leaf_rope))
    return hipErrorOutOfMemory;

  // ---- scene box
  Region phase("ArborX::BVH::BVH::calculate_scene_bounding_box");
  if (geometry == GEOM_POINTS)
    AXB_LAUNCH(k_scene_partial<GEOM_POINTS>, dim3(partial_blocks), dim3(256), 0,
                       stream, values, n, partials);
  else
    AXB_LAUNCH(k_scene_partial<GEOM_BOXES>, dim3(partial_blocks), dim3(256), 0,
                       stream, values, n, partials);
  AXB_LAUNCH(k_scene_final, dim3(1), dim3(256), 0, stream, partials, partial_blocks,
             out.scene);

  // ---- Morton codes + digit histograms
  phase.next("ArborX::BVH::BVH::compute_linear_ordering");
  e0 = fill_u32(hist, 0u, 8 * 256, stream);
  if (e0 != hipSuccess)
    return e0;
  {
    const unsigned g = (unsigned)std::min<int64_t>(2048, div_up(n, 256));
    if (geometry == GEOM_POINTS)
      AXB_LAUNCH(k_morton_hist<GEOM_POINTS>, dim3(g), dim3(256), 0, stream, values,
                         n, out.scene, keys0, hist);
    else
      AXB_LAUNCH(k_morton_hist<GEOM_BOXES>, dim3(g), dim3(256), 0, stream, values,
                         n, out.scene, keys0, hist);
  }

  // ---- sort (8 passes of 8 bits; result back in keys0 / vals0)
  phase.next("ArborX::BVH::BVH::sort_linearized_order");
  hipError_t e = radix_sort_pairs_u64(keys0, keys1, vals0, vals1, n, 8, hist, ws, err, stream);
  if (e != hipSuccess)
    return e;

  // ---- hierarchy
  phase.next("ArborX::BVH::BVH::generate_hierarchy");
  e = fill_u32((uint32_t *)c, 0u, n, stream);
  if (e != hipSuccess)
    return e;
  {
    const unsigned g = (unsigned)div_up(n, CLIMB_LEAVES);
    const int npend = 2 * (int)g;
    const unsigned gp = grid_for(npend);
#define AXB_CLIMB(GEO, ROPE)                                                                  \
  do                                                                                          \
  {                                                                                           \
    AXB_LAUNCH((k_climb_local<GEO, ROPE>), dim3(g), dim3(CLIMB_LEAVES), 0, stream, values,    \
               ni, keys0, vals0, ranges, inodes, leafrec, leaf_rope, c, top, pending,         \
               spine_count, spine_list);                                                      \
    AXB_LAUNCH((k_climb_global<GEO, ROPE>), dim3(gp), dim3(256), 0, stream, pending, npend,   \
               ni, keys0, ranges, inodes, leafrec, c, top, spine_count, spine_list);          \
  } while (0)
    if (geometry == GEOM_POINTS)
    {
      if (keep_topology)
        AXB_CLIMB(GEOM_POINTS, true);
      else
        AXB_CLIMB(GEOM_POINTS, false);
    }
    else
    {
      if (keep_topology)
        AXB_CLIMB(GEOM_BOXES, true);
      else
        AXB_CLIMB(GEOM_BOXES, false);
    }
#undef AXB_CLIMB
  }
  e = inclusive_scan_i32(c, out.C, n, ws, err, stream);
  if (e != hipSuccess)
    return e;

  // ---- pre-order linearisation
  {
    const unsigned gi = grid_for(nint), gl = grid_for(n);
    if (node_format == NODES_FP32)
    {
      if (geometry == GEOM_POINTS)
      {
        AXB_LAUNCH(k_lin_fp32<GEOM_POINTS>, dim3(gi), dim3(256), 0, stream, inodes,
                           nint, out.C, out.slots, (uint32_t)out.max_slots, err);
        AXB_LAUNCH((k_lin_leaves<GEOM_POINTS, 2>), dim3(gl), dim3(256), 0, stream,
                           leafrec, ni, out.C, out.slots, (uint32_t)out.max_slots, err);
      }
      else
      {
        AXB_LAUNCH(k_lin_fp32<GEOM_BOXES>, dim3(gi), dim3(256), 0, stream, inodes,
                           nint, out.C, out.slots, (uint32_t)out.max_slots, err);
        AXB_LAUNCH((k_lin_leaves<GEOM_BOXES, 2>), dim3(gl), dim3(256), 0, stream,
                           leafrec, ni, out.C, out.slots, (uint32_t)out.max_slots, err);
      }
    }
    else
    {
      if (geometry == GEOM_POINTS)
      {
        AXB_LAUNCH(k_lin_q16<GEOM_POINTS>, dim3(grid_for(nlists, 4)), dim3(256), 0, stream,
                   inodes, ni, spine_count, spine_list, (int)nlists, top, out.C, out.scene,
                   out.slots, (uint32_t)out.max_slots, err);
        AXB_LAUNCH((k_lin_leaves<GEOM_POINTS, 1>), dim3(gl), dim3(256), 0, stream,
                           leafrec, ni, out.C, out.slots, (uint32_t)out.max_slots, err);
      }
      else
      {
        AXB_LAUNCH(k_lin_q16<GEOM_BOXES>, dim3(grid_for(nlists, 4)), dim3(256), 0, stream,
                   inodes, ni, spine_count, spine_list, (int)nlists, top, out.C, out.scene,
                   out.slots, (uint32_t)out.max_slots, err);
        AXB_LAUNCH((k_lin_leaves<GEOM_BOXES, 1>), dim3(gl), dim3(256), 0, stream,
                           leafrec, ni, out.C, out.slots, (uint32_t)out.max_slots, err);
      }
    }
  }
  if (geometry == GEOM_POINTS)
  {
    if (node_format == NODES_FP32)
      AXB_LAUNCH((k_pad_tail<1, 2>), dim3(1), dim3(64), 0, stream, out.slots, out.C, ni);
    else
      AXB_LAUNCH((k_pad_tail<1, 1>), dim3(1), dim3(64), 0, stream, out.slots, out.C, ni);
  }
  else
  {
    if (node_format == NODES_FP32)
      AXB_LAUNCH((k_pad_tail<2, 2>), dim3(1), dim3(64), 0, stream, out.slots, out.C, ni);
    else
      AXB_LAUNCH((k_pad_tail<2, 1>), dim3(1), dim3(64), 0, stream, out.slots, out.C, ni);
  }
  e = hipGetLastError();

  dfree(keys1, stream);
  dfree(vals1, stream);
  dfree(hist, stream);
  dfree(partials, stream);
  dfree(ws, stream);
  dfree(ranges, stream);
  dfree(leafrec, stream);
  dfree(c, stream);
  dfree(top, stream);
  dfree(pending, stream);
  dfree(spine_count, stream);
  dfree(spine_list, stream);
  if (keep_topology)
  {
    out.keys = keys0;
    out.perm = vals0;
    out.leaf_rope = leaf_rope;
    out.inodes = inodes;
  }
  else
  {
    dfree(keys0, stream);
    dfree(vals0, stream);
    dfree(inodes, stream);
  }
  return e;
}

hipError_t build_status(const BuildOut &b, int64_t n, uint32_t *status, hipStream_t stream)
{
  AXB_LAUNCH(k_build_status, dim3(1), dim3(64), 0, stream, b.scene, b.C, n, status);
  return hipGetLastError();
}

hipError_t export_topology(const INode *inodes, int64_t nint, int32_t *left_child,
                           int32_t *rope, float *boxes, hipStream_t stream)
{
  if (nint <= 0)
    return hipSuccess;
  AXB_LAUNCH(k_export_topology, dim3(grid_for(nint)), dim3(256), 0, stream, inodes,
                     (int)nint, left_child, rope, boxes);
  return hipGetLastError();
}

} // namespace axb
