// axb_query.hip -- spatial and nearest traversals on gfx950.
//
// Spatial: the reference's stackless rope walk (detail/ArborX_TreeTraversal.hpp:
// 97-119) over the pre-order slot array: an internal node that passes its
// test continues at the next record (its left child), one that fails jumps
// to `slot + skip` (its rope); a leaf is tested and the walk moves to the
// next record.  One lane per predicate, predicates in Morton-32 order.
// Sphere tests use `d2 <= sqrt_threshold(r)`, which decides exactly the
// reference's `sqrt(d2) <= r` (Intersects.hpp:82-114).
//
// Nearest: result = the k smallest (distance, index) pairs, distance =
// RN(sqrt(d2)) as in Distance.hpp:52-80.  k_knn_collect, one lane per query:
//   1. greedy descent to the leaf nearest in tree order; the k-th distance
//      R of the W = 6k leaves around it in Morton order (contiguous in
//      pre-order) bounds the true k-th distance from above (they are real
//      points);
//   2. a fixed-radius rope walk appending (d2, index) of every leaf within R
//      to a row of `cap` candidates; a full row shrinks R in place and the
//      walk goes on; a few rounds adapt R until the row holds between k and
//      cap leaves;
//   3. exact top-k of the row, (d2, index) keys sorted and merged by
//      bitonic networks in registers.
// Queries left unresolved go to k_knn_fallback (exact adaptive walk with an
// LDS candidate buffer); k > 64 uses a binary max-heap in the output row
// (the reference's NearestBufferProvider scheme, NearestBufferProvider.hpp:
// 23-73).
#include "axb_internal.hpp"

#include <algorithm>

namespace axb
{

namespace
{

// One lane per predicate (in `order` when given): rope traversal of the
// whole tree; row q of the CRS (counts, scratch rows or the final rows).
// COUNT: the launch adds its visits to the visit counters (profiling
// passes); without it the walk carries no counters.
template <int F, int G, int PK, int MODE, bool SLOTS = false, bool COUNT = true>
__global__ __launch_bounds__(256) void k_spatial(
    TreeView tv, const float *__restrict__ preds, int64_t nq, float radius,
    const uint32_t *__restrict__ order, int64_t *__restrict__ offsets,
    int32_t *__restrict__ out, int32_t buffer)
{
  const int64_t t = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= nq)
    return;
  const int64_t q = order ? (int64_t)order[t] : t;
  SpatialPred pr;
  if (PK == PRED_RAYS)
    pr = ray_spatial_pred(preds + 6 * q);
  else if (PK == PRED_BOXES)
  {
    const float *p = preds + 6 * q;
    pr = box_pred(Box6{{p[0], p[1], p[2]}, {p[3], p[4], p[5]}});
  }
  else
  {
    const float *p = preds + (PK == PRED_SPHERES ? 4 : 3) * q;
    pr = sphere_pred(p[0], p[1], p[2], PK == PRED_SPHERES ? p[3] : radius);
  }
  uint32_t v_int = 0, v_leaf = 0;
  if constexpr (MODE != MODE_COUNT)
  {
    __shared__ int32_t s_stage[ROW_STAGE * 256];
    StagedRow<MODE> rw;
    rw.init(out, offsets, q, buffer, s_stage + threadIdx.x);
    rope_walk<F, G, PK, SLOTS>(tv, pr, 0u, tv.total, ~0u, rw, v_int, v_leaf);
    row_epilogue(tv, rw, offsets, q, COUNT ? v_int : 0u, COUNT ? v_leaf : 0u);
  }
  else
  {
    RowWriter<MODE> rw;
    rw.init(out, offsets, q, buffer);
    rope_walk<F, G, PK, SLOTS>(tv, pr, 0u, tv.total, ~0u, rw, v_int, v_leaf);
    row_epilogue(tv, rw, offsets, q, COUNT ? v_int : 0u, COUNT ? v_leaf : 0u);
  }
}

// Rows of leaf slots hit by each ray -> rows of values in (distance, index)
// order, one lane per row, one launch per row-length class so that the
// short rows keep few registers and many wavefronts in flight: CLASS 0
// rows of up to 16, 1 up to 32, 2 the longer ones.
template <int CLASS>
__global__ __launch_bounds__(256) void k_ray_sort_rows(TreeView tv, const float *__restrict__ rays,
                                                       int64_t nq,
                                                       const int64_t *__restrict__ offsets,
                                                       int32_t *__restrict__ rows)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq)
    return;
  const int64_t b = offsets[q], m = offsets[q + 1] - b;
  const int cls = m <= 16 ? 0 : (m <= 32 ? 1 : 2);
  if (m == 0 || cls != CLASS)
    return;
  const RayPred ray = ray_pred(rays + 6 * q);
  if (CLASS == 0)
  {
    if (m <= 8)
      sort_ray_row_n<8>(tv, ray, rows + b, (int)m);
    else
      sort_ray_row_n<16>(tv, ray, rows + b, (int)m);
  }
  else if (CLASS == 1)
    sort_ray_row_n<32>(tv, ray, rows + b, (int)m);
  else
    sort_ray_row(tv, ray, rows + b, m);
}

// Whole rows of an ordered query with a sphere, box or point
// (OrderedSpatial<G>, every leaf at a finite distance): row q holds the
// slots of all n leaves (leaf j at S_leaf * j + S_int * C[j]) ...
template <int F, int G>
__global__ __launch_bounds__(256) void k_leaf_slot_rows(TreeView tv, int64_t nq,
                                                        int32_t *__restrict__ rows)
{
  constexpr uint32_t SL = leaf_slots<G>(), SI = internal_slots<F>();
  const int64_t n = tv.n;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq * n)
    return;
  const int64_t j = i % n;
  rows[i] = (int32_t)(SL * (uint32_t)j + SI * (uint32_t)tv.C[j]);
}
// ... then sorted by (distance(G, leaf), index) and turned into values, one
// lane per row.
template <int G, int PK>
__global__ __launch_bounds__(256) void k_ordered_sort_rows(TreeView tv,
                                                           const float *__restrict__ preds,
                                                           int64_t nq,
                                                           int32_t *__restrict__ rows)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq)
    return;
  constexpr int stride = PK == PRED_SPHERES ? 4 : (PK == PRED_BOXES ? 6 : 3);
  const OrderedGeom g = ordered_geom(PK, preds + stride * q);
  const int64_t n = tv.n;
  sort_ray_row<G>(tv, g, rows + q * n, n);
}

__global__ __launch_bounds__(256) void k_row_prefix_counts(const int64_t *__restrict__ offsets,
                                                           int64_t nq, int32_t k,
                                                           int64_t *__restrict__ counts)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q > nq)
    return;
  const int64_t m = q < nq ? offsets[q + 1] - offsets[q] : 0;
  counts[q] = m < (int64_t)k ? m : (int64_t)k;
}
__global__ __launch_bounds__(256) void k_row_prefix_copy(const int64_t *__restrict__ offsets,
                                                         const int32_t *__restrict__ rows,
                                                         int64_t nq,
                                                         const int64_t *__restrict__ out_offsets,
                                                         int32_t *__restrict__ out)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq)
    return;
  const int64_t a = offsets[q], b = out_offsets[q], m = out_offsets[q + 1] - b;
  for (int64_t j = 0; j < m; ++j)
    out[b + j] = rows[a + j];
}

// Row writer that stops the walk after `limit` hits.
template <int MODE>
struct LimitedRow
{
  RowWriter<MODE> *rw;
  int64_t limit;
  AXB_D bool hit(int32_t v)
  {
    rw->hit(v);
    return rw->cnt < limit;
  }
};

// Ordered ray queries (OrderedSpatial<Ray>, and Nearest<Ray> with limit =
// k): one lane per ray, the best-first walk, row q in (distance, index)
// order.  The walk's heap lives in private (scratch) memory.
template <int F, int MODE>
__global__ __launch_bounds__(256) void k_ordered(TreeView tv, const float *__restrict__ rays,
                                                 int64_t nq, int32_t limit,
                                                 const uint32_t *__restrict__ order,
                                                 int64_t *__restrict__ offsets,
                                                 int32_t *__restrict__ out, int32_t buffer)
{
  const int64_t t = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= nq)
    return;
  const int64_t q = order ? (int64_t)order[t] : t;
  const RayPred ray = ray_pred(rays + 6 * q);
  RowWriter<MODE> rw;
  rw.init(out, offsets, q, buffer);
  LimitedRow<MODE> sink{&rw, (int64_t)limit};
  uint32_t v_int = 0, v_leaf = 0;
  if (limit > 0)
    ordered_walk<F>(tv, ray, sink, v_int, v_leaf);
  row_epilogue(tv, rw, offsets, q, v_int, v_leaf);
}

// Ray construction (ArborX_Ray.hpp:42-51, misc/ArborX_Vector.hpp:41-55 and
// :77-87): the norm is the square root of the double sum r = 0;
// r += v_d * v_d, each component v_d / norm in double rounded to float, a
// zero component stored as +0.
__global__ __launch_bounds__(256) void k_make_rays(const float *__restrict__ origins,
                                                   const float *__restrict__ dirs, int64_t n,
                                                   float *__restrict__ rays,
                                                   uint32_t *__restrict__ err)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  double r = 0.0;
  double v[3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    v[d] = (double)dirs[3 * i + d];
    r = r + v[d] * v[d];
  }
  const double mag = __builtin_sqrt(r);
  if (!(mag > 0.0))
    raise_error(err, KERR_LAYOUT);
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    rays[6 * i + d] = origins[3 * i + d];
    const float w = (float)(v[d] / mag);
    rays[6 * i + 3 + d] = w == 0.f ? 0.f : w;
  }
}

// Neighbour lists over a point tree (NeighborList.hpp:49-188 semantics):
// lane t takes the leaf at position L = list ? list[t] : t and the sphere
// {its point, radius}; row ~w (its value) receives the values of the other
// leaves in the sphere -- all of them (FULL) or only those whose leaf comes
// before L (half list: HalfTraversal.hpp pairs (earlier, later) land in the
// later leaf's row).  The half walk covers the slots before L's record,
// which every ancestor of L precedes and passes (distance 0).  rowof[L] is
// written in the first pass (leaf position -> row) for the overflow list.
template <int F, bool FULL, int MODE>
__global__ __launch_bounds__(256) void k_nlist(TreeView tv, float radius, int64_t nl,
                                               const uint32_t *__restrict__ list,
                                               int64_t *__restrict__ offsets,
                                               int32_t *__restrict__ out, int32_t buffer,
                                               uint32_t *__restrict__ rowof)
{
  constexpr uint32_t SL = leaf_slots<GEOM_POINTS>(), SI = internal_slots<F>();
  const int64_t t = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= nl)
    return;
  const uint32_t L = list ? list[t] : (uint32_t)t;
  const uint32_t self = SL * L + SI * (uint32_t)tv.C[L];
  const uint4 s = tv.slots[self];
  const int64_t q = ~(int32_t)s.w;
  if (rowof)
    rowof[L] = (uint32_t)q;
  const SpatialPred pr = sphere_pred(bits_f(s.x), bits_f(s.y), bits_f(s.z), radius);
  RowWriter<MODE> rw;
  rw.init(out, offsets, q, buffer);
  uint32_t v_int = 0, v_leaf = 0;
  rope_walk<F, GEOM_POINTS, PRED_SPHERES>(tv, pr, 0u, FULL ? tv.total : self, self, rw, v_int,
                                          v_leaf);
  row_epilogue(tv, rw, offsets, q, v_int, v_leaf);
}

// ------------------------------------------------------------ nearest

// The K smallest values seen (float bits, ascending; no indices): the seed
// only needs the k-th distance, in half the registers of RegList.  It is fed
// squared distances: RN(sqrt) is monotone, so the k-th smallest distance is
// RN(sqrt) of the k-th smallest squared distance, one square root per query.
template <int K>
struct DistList
{
  uint32_t a[K];
  // The first K - k entries hold 0, below every value inserted (all >= 0):
  // they stay in front, and the k-th smallest value inserted is always
  // a[K - 1], read at a fixed index.  (Picking a[k - 1] for a run-time k
  // turns into an indexed read, which keeps the list in scratch memory for
  // K > 16.)
  AXB_D void init(int k)
  {
#pragma unroll
    for (int j = 0; j < K; ++j)
      a[j] = j < K - k ? 0u : ~0u;
  }
  AXB_D void insert(float d)
  {
    const uint32_t x = f_bits(d); // d >= 0: bit order = value order
#pragma unroll
    for (int j = K - 1; j > 0; --j)
    {
      const uint32_t hi = a[j - 1] < x ? x : a[j - 1];
      a[j] = hi < a[j] ? hi : a[j];
    }
    a[0] = x < a[0] ? x : a[0];
  }
  // RN(sqrt) of the k-th smallest (init's k, 1 <= k <= K); +inf while fewer
  // than k were inserted
  AXB_D float kth() const
  {
    const uint32_t v = a[K - 1];
    return v == ~0u ? INFINITY : sqrt_rn(bits_f(v));
  }
};


template <int G>
AXB_D float leaf_d2(float qx, float qy, float qz, uint4 s0, uint4 s1)
{
  const float x = bits_f(s0.x), y = bits_f(s0.y), z = bits_f(s0.z);
  if (G == GEOM_POINTS)
    return dist2_pp(qx, qy, qz, x, y, z);
  const Box6 lb = {{x, y, z}, {bits_f(s1.x), bits_f(s1.y), bits_f(s1.z)}};
  return dist2_pnode(qx, qy, qz, lb);
}

// Rope walk over [p, pend) for nearest search: a node farther than the
// current k-th distance is skipped (`d2 > thr`, ties kept), a leaf closer
// than it is inserted.  `rank` is the Morton rank of the next leaf; leaves
// of rank [j0, j0 + m) were seeded already.  The subtree at skip_at (already
// walked) is jumped over.
// Candidates (leaves within the current k-th distance) wait in a per-lane
// LDS buffer; the list is updated in flushes that the whole wavefront takes
// together once some lane's buffer is full, so the insertion network runs a
// few times per wave instead of at almost every step of a divergent walk.
constexpr int KNN_BUF = 8;

struct CandBuf
{
  float *d2;   // KNN_BUF x blockDim, strided by lane
  int32_t *id;
  int stride, lane, cnt;
};

template <class List>
AXB_D void knn_flush(List &L, float &thr, CandBuf &B)
{
#pragma unroll 1
  for (int j = 0; j < KNN_BUF; ++j)
  {
    if (j < B.cnt)
    {
      const float d2 = B.d2[j * B.stride + B.lane];
      if (d2 <= thr)
      {
        const float dist = sqrt_rn(d2);
        const int32_t idx = B.id[j * B.stride + B.lane];
        if (L.accepts(dist, idx))
        {
          L.insert(dist, idx);
          if (L.full())
            thr = sqrt_threshold(L.kth_d);
        }
      }
    }
  }
  B.cnt = 0;
}

// Greedy descent (closer child first, ties left) from the root to a leaf
// record near the query.  A left child that is a leaf starts a run of at
// most FLAT_MAX leaves (a subtree that is not stored); the run is scanned
// to find the right child, the first internal record inside the parent's
// subtree after it (if the right child is a run as well, the parent holds
// only leaves and the nearest of them ends the descent).  Returns the slot
// of the leaf reached.  The deepest stored node on the path with at least
// `min_slots` slots is returned in X (slot, slots, box).
struct Subtree
{
  uint32_t p, slots;
  Box6 box;
};
template <int F, int G>
AXB_D uint32_t greedy_descent(const TreeView &tv, float qx, float qy, float qz,
                              uint32_t min_slots, Subtree &X, uint32_t &v_int,
                              uint32_t &v_leaf)
{
  constexpr uint32_t SL = leaf_slots<G>(), SI = internal_slots<F>();
  const uint32_t end = tv.total;
  uint32_t p = 0;
  Box6 cur = tv.scene;
  float cs[3] = {tv.ss[0], tv.ss[1], tv.ss[2]};
  X.p = 0;
  X.slots = end;
  {
    const uint4 s0 = tv.slots[0], s1 = tv.slots[1];
    if ((int32_t)s0.w >= 0)
    {
      ++v_int;
      node_box<F>(tv, s0, s1, false, cur, cs, cur);
      set_frame<F>(cur, cur, cs);
    }
    X.box = cur;
  }
  while (true)
  {
    const uint4 s0 = tv.slots[p];
    const int32_t w = (int32_t)s0.w;
    if (w < 0)
      return p;
    const uint32_t sub_end = p + (uint32_t)w; // end of this node's subtree
    const uint32_t lp = p + SI;
    const uint4 l0 = tv.slots[lp], l1 = tv.slots[lp + 1];
    const int32_t lw = (int32_t)l0.w;
    Box6 lb;
    float dl;
    uint32_t rp, lbest = lp;
    if (lw < 0)
    {
      // the left child is a run of leaves: nearest of them, and the slot
      // after the run
      dl = INFINITY;
      rp = lp;
      for (int j = 0; j < 2 * FLAT_MAX && rp < sub_end; ++j)
      {
        const uint4 a0 = tv.slots[rp];
        if ((int32_t)a0.w >= 0)
          break;
        ++v_leaf;
        const float d = leaf_d2<G>(qx, qy, qz, a0, SL == 2 ? tv.slots[rp + 1] : a0);
        if (d < dl)
          dl = d, lbest = rp;
        rp += SL;
      }
      if (rp >= sub_end || (int32_t)tv.slots[rp].w < 0)
        return lbest; // only leaves below this node
    }
    else
    {
      ++v_int;
      node_box<F>(tv, l0, l1, true, cur, cs, lb);
      dl = dist2_pnode(qx, qy, qz, lb);
      rp = lp + (uint32_t)lw;
    }
    if (rp >= end)
    {
      raise_error(tv.err, KERR_LAYOUT);
      return lbest;
    }
    const uint4 r0 = tv.slots[rp], r1 = tv.slots[rp + 1];
    const int32_t rw = (int32_t)r0.w;
    Box6 rb;
    float dr;
    if (rw < 0)
    {
      ++v_leaf;
      dr = leaf_d2<G>(qx, qy, qz, r0, r1);
    }
    else
    {
      ++v_int;
      node_box<F>(tv, r0, r1, false, cur, cs, rb);
      dr = dist2_pnode(qx, qy, qz, rb);
    }
    if (dl <= dr)
    {
      if (lw < 0)
        return lbest;
      p = lp;
      set_frame<F>(lb, cur, cs);
      if ((uint32_t)lw >= min_slots)
        X = Subtree{lp, (uint32_t)lw, lb};
    }
    else
    {
      p = rp;
      if (rw < 0)
        return rp;
      set_frame<F>(rb, cur, cs);
      if ((uint32_t)rw >= min_slots)
        X = Subtree{rp, (uint32_t)rw, rb};
    }
  }
}

// The W leaves nearest to slot p in slot order (about half before it),
// fn(s0, s1) for each; [ws, we) is the slot range they span (no other leaf
// lies in it).  W must not exceed the number of leaves.  (One slot per
// step: fn is inlined at three sites only, which keeps a sorted list of
// up to 64 entries in registers.)
template <int G, class Fn>
AXB_D void slot_window(const TreeView &tv, uint32_t p, int W, uint32_t &ws, uint32_t &we,
                       Fn &&fn, uint32_t &v_int, uint32_t &v_leaf)
{
  constexpr uint32_t SL = leaf_slots<G>();
  const uint32_t end = tv.total;
  int got = 0;
  uint32_t s = p, e = p;
  while (got < W / 2 && s > 0)
  {
    --s;
    const uint4 s0 = tv.slots[s];
    if ((int32_t)s0.w < 0)
    {
      ++v_leaf;
      fn(s0, SL == 2 ? tv.slots[s + 1] : s0);
      ++got;
    }
    else
      ++v_int;
  }
  while (got < W && e < end)
  {
    const uint4 s0 = tv.slots[e];
    if ((int32_t)s0.w < 0)
    {
      ++v_leaf;
      fn(s0, SL == 2 ? tv.slots[e + 1] : s0);
      ++got;
      e += SL;
    }
    else
    {
      ++v_int;
      ++e;
    }
  }
  while (got < W && s > 0) // near the end of the array: more before p
  {
    --s;
    const uint4 s0 = tv.slots[s];
    if ((int32_t)s0.w < 0)
    {
      ++v_leaf;
      fn(s0, SL == 2 ? tv.slots[s + 1] : s0);
      ++got;
    }
    else
      ++v_int;
  }
  ws = s;
  we = e;
}

template <int F, int G, class List>
AXB_D void knn_walk(const TreeView &tv, float qx, float qy, float qz, List &L, float &thr,
                    CandBuf &B, uint32_t p, uint32_t pend, bool desc, Box6 par, float ps[3],
                    uint32_t skip_at, uint32_t skip_slots, uint32_t ws, uint32_t we,
                    uint32_t &v_int, uint32_t &v_leaf)
{
  constexpr uint32_t SL = leaf_slots<G>(), SI = internal_slots<F>();
  uint4 s0 = make_uint4(0, 0, 0, 0), s1 = s0;
  // Outer loop: the whole wavefront; inner loops ("while-while"): a run of
  // internal nodes up to the next leaf, then a run of leaves up to the next
  // internal node (or a full candidate buffer).  Lanes that finish an inner
  // run early wait instead of executing the other record kind's code.
  while (__ballot(p < pend))
  {
    if (__ballot(B.cnt == KNN_BUF))
      knn_flush(L, thr, B);
    while (p < pend)
    {
      if (p == skip_at)
      {
        p += skip_slots;
        desc = false;
        continue;
      }
      s0 = tv.slots[p];
      s1 = (F == NODES_FP32 || G == GEOM_BOXES) ? tv.slots[p + 1] : s0;
      const int32_t w = (int32_t)s0.w;
      if (w < 0)
        break;
      ++v_int;
      Box6 b;
      node_box<F>(tv, s0, s1, desc, par, ps, b);
      if (dist2_pnode(qx, qy, qz, b) <= thr)
      {
        set_frame<F>(b, par, ps);
        desc = true;
        p += SI;
      }
      else
      {
        if (w == 0)
        {
          raise_error(tv.err, KERR_LAYOUT);
          p = pend;
          break;
        }
        desc = false;
        p += (uint32_t)w;
      }
    }
    while (p < pend && (int32_t)s0.w < 0 && B.cnt < KNN_BUF)
    {
      ++v_leaf;
      if (p < ws || p >= we) // the seed window's leaves are in the list already
      {
        const float d2 = leaf_d2<G>(qx, qy, qz, s0, s1);
        if (d2 <= thr)
        {
          B.d2[B.cnt * B.stride + B.lane] = d2;
          B.id[B.cnt * B.stride + B.lane] = ~(int32_t)s0.w;
          ++B.cnt;
        }
      }
      p += SL;
      desc = false;
      if (p >= pend || p == skip_at)
        break;
      s0 = tv.slots[p];
      s1 = G == GEOM_BOXES ? tv.slots[p + 1] : s0;
    }
  }
  knn_flush(L, thr, B);
}

template <int F, int G, class List>
AXB_D void knn_traverse(const TreeView &tv, float qx, float qy, float qz, int m, int W,
                        List &L, CandBuf &B, uint32_t &v_int, uint32_t &v_leaf)
{
  constexpr uint32_t SL = leaf_slots<G>(), SI = internal_slots<F>();
  const uint32_t end = tv.total;
  const int n = tv.n;
  // ---- phase 1: greedy descent to a leaf near the query, keeping the
  // deepest node X on the path with about >= lmin leaves (by its slots)
  const int lmin = 4 * m > 32 ? 4 * m : 32;
  Subtree X;
  const uint32_t p = greedy_descent<F, G>(tv, qx, qy, qz, (uint32_t)lmin * (SL + SI), X, v_int,
                                          v_leaf);
  // ---- seed: the W >= m leaves around it in slot order
  uint32_t ws = 0, we = 0;
  slot_window<G>(
      tv, p, W, ws, we,
      [&](uint4 s0, uint4 s1) __attribute__((always_inline)) {
        const float dist = sqrt_rn(leaf_d2<G>(qx, qy, qz, s0, s1));
        const int32_t idx = ~(int32_t)s0.w;
        if (L.accepts(dist, idx))
          L.insert(dist, idx);
      },
      v_int, v_leaf);
  if (W >= n)
    return; // every leaf is already in the list
  float thr = L.full() ? sqrt_threshold(L.kth_d) : INFINITY;
  float ps[3];
  Box6 DX = X.box;
  set_frame<F>(DX, DX, ps);
  // ---- phase 2: the local subtree X first (tightens the k-th distance),
  // then the whole tree with X jumped over.  One walk per segment.
  const bool local = X.slots < end;
  for (int seg = local ? 0 : 1; seg < 2; ++seg)
  {
    const bool first_seg = seg == 0;
    knn_walk<F, G>(tv, qx, qy, qz, L, thr, B, first_seg ? X.p + SI : 0u,
                   first_seg ? X.p + X.slots : end, first_seg, DX, ps,
                   (local && !first_seg) ? X.p : UINT32_MAX, X.slots, ws, we, v_int, v_leaf);
  }
}

// ------------------------------------------------ nearest: collect + select
// For k <= 64 the search runs as two kernels:
//  collect  greedy descent, then the k-th distance R0 of the W = 6k leaves
//           around the query in Morton order (an upper bound of the true
//           k-th distance: those are k real points within R0); then a
//           fixed-radius rope walk that appends every leaf with
//           RN(sqrt(d2)) <= R to the query's candidate row (cap entries),
//           shrinking R in place when the row fills; then the exact
//           (distance, index) top-k of the row;
//  fallback the few queries left unresolved after the rounds run the exact
//           adaptive walk (knn_traverse), once after all batches.
// The walk of `collect` carries no list, so it runs at the register budget
// of the spatial traversal.

// Index of slot s (< 256) of a wavefront's seed window in its block's
// candidate stage seen as 1024 uint4: the bytes of the stage entries of the
// wavefront's own 64 lanes (entry j of lane l at uint2 j * 256 + l), 32
// slots in each of the 8 entry rows, so that no other wavefront's entries
// are touched.
AXB_D uint32_t seed_window_index(uint32_t s, uint32_t wave)
{
  return ((s >> 5) << 7) + (wave << 5) + (s & 31u);
}

// Seed radius of a query: the k-th smallest distance over the leaves of a
// window of about W leaves around the leaf its greedy descent reaches (an
// upper bound of the k-th nearest distance); +inf when the window holds
// fewer than k leaves.  stage: the block's candidate stage (KNN_STAGE x 256
// uint2, free during the seed) or null.
template <int F, int G, int K>
AXB_D float knn_seed(const TreeView &tv, float qx, float qy, float qz, int k, int W,
                     uint32_t &v_int, uint32_t &v_leaf, uint2 *stage)
{
  Subtree X;
  const uint32_t p = greedy_descent<F, G>(tv, qx, qy, qz, UINT32_MAX, X, v_int, v_leaf);
  if constexpr (G == GEOM_POINTS)
  {
    // The slots [a, a + S) around p, S = 1.5 W rounded up to 8 (about W
    // leaves), NB at a time: keys = d2 bits of the leaf records, ~0 for the
    // others; each batch is sorted and merged into the K smallest by
    // bitonic networks on 32-bit keys (min / max pairs), all loads of a
    // batch in flight together.  The first K - k entries hold 0 so that the
    // k-th smallest is A[K - 1].
    constexpr int NB = K < 8 ? K : 8;
    const uint32_t total = tv.total;
    const uint32_t S = (uint32_t)((3 * W / 2 + 7) & ~7);
    uint32_t a = p > S / 2 ? p - S / 2 : 0u;
    if (a + S > total)
      a = total > S ? total - S : 0u;
    const uint32_t e = a + S < total ? a + S : total;
    uint32_t A[K];
#pragma unroll
    for (int j = 0; j < K; ++j)
      A[j] = j < K - k ? 0u : ~0u;
    uint32_t leaves = 0;
    // The windows of a wavefront's queries (consecutive in Morton order)
    // overlap: when the wavefront is full and the union of its windows has
    // at most 256 slots, the wavefront loads the union once into LDS (every
    // load in flight together) and each lane scans its window there.
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint32_t ua = a, ue = e;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
    {
      const uint32_t oa = (uint32_t)__shfl_xor((int)ua, o, 64);
      const uint32_t oe = (uint32_t)__shfl_xor((int)ue, o, 64);
      ua = oa < ua ? oa : ua;
      ue = oe > ue ? oe : ue;
    }
    uint4 *win = reinterpret_cast<uint4 *>(stage);
    const bool coop = stage != nullptr && __ballot(1) == ~0ull && ue - ua <= 256u;
    if (coop)
    {
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j)
      {
        const uint32_t sl = j * 64u + lane;
        if (ua + sl < ue)
          win[seed_window_index(sl, wv)] = tv.slots[ua + sl];
      }
      // the wavefront's LDS writes complete before any lane reads
      __builtin_amdgcn_s_waitcnt(0xc07f); // lgkmcnt(0)
      __builtin_amdgcn_wave_barrier();
    }
    for (uint32_t b = a; b < e; b += NB)
    {
      uint32_t B[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j)
      {
        const uint32_t slot = b + (uint32_t)j;
        const uint32_t sc = slot < e ? slot : e - 1;
        const uint4 r = coop ? win[seed_window_index(sc - ua, wv)] : tv.slots[sc];
        const bool leaf = slot < e && (int32_t)r.w < 0;
        B[j] = leaf ? f_bits(dist2_pp(qx, qy, qz, bits_f(r.x), bits_f(r.y), bits_f(r.z)))
                    : ~0u;
        leaves += leaf;
      }
      bitonic_sort<NB>(B);
#pragma unroll
      for (int j = K - NB; j < K; ++j)
        A[j] = A[j] < B[K - 1 - j] ? A[j] : B[K - 1 - j];
      bitonic_merge<K>(A);
    }
    v_leaf += leaves;
    v_int += (e - a) - leaves;
    const int m = k < tv.n ? k : tv.n;
    if ((int)leaves >= m) // else the leaf-counting window below
      return A[K - 1] == ~0u ? INFINITY : sqrt_rn(bits_f(A[K - 1]));
  }
  DistList<K> L;
  L.init(k);
  uint32_t ws, we;
  slot_window<G>(
      tv, p, W, ws, we,
      [&](uint4 s0, uint4 s1) __attribute__((always_inline)) { L.insert(leaf_d2<G>(qx, qy, qz, s0, s1)); }, v_int,
      v_leaf);
  return L.kth();
}

// Fixed-radius rope walk appending every leaf with d2 <= thr to `row` as
// {index, d2 bits}: read as one little-endian 64-bit word that is the
// (d2, index) key, d2 bits high (d2 >= 0, so bit order is value order).  Stops as soon as more than `cap` leaves qualify;
// returns the count (cap + 1 on overflow).
// Candidates of a point walk gather in LDS, KNN_STAGE per lane (entry j of
// lane l at st[j * 256 + l]), and leave for the row as whole 64-byte
// chunks: each chunk of the row is written once, complete (a row's lanes
// write far apart, so single 8-byte stores left most of each 128-byte line
// to be written back partially, several times).
constexpr int KNN_STAGE = 8;
struct CandStage
{
  uint2 *st;
  int lane;
  AXB_D void put(int c, uint2 e) { st[(c & (KNN_STAGE - 1)) * 256 + lane] = e; }
  AXB_D uint2 get(int c) const { return st[(c & (KNN_STAGE - 1)) * 256 + lane]; }
  // the chunk holding entries [c0, c0 + 8) of the row, c0 a multiple of 8
  AXB_D void flush(uint2 *__restrict__ cand, uint32_t row_at, int c0) const
  {
    uint4 *dst = reinterpret_cast<uint4 *>(cand + row_at + c0);
#pragma unroll
    for (int j = 0; j < KNN_STAGE / 2; ++j)
    {
      const uint2 a = st[(2 * j) * 256 + lane], b = st[(2 * j + 1) * 256 + lane];
      dst[j] = make_uint4(a.x, a.y, b.x, b.y);
    }
  }
  // reloads the partial chunk of a row of c entries from the row
  AXB_D void reload(const uint2 *__restrict__ cand, uint32_t row_at, int c)
  {
    const int c0 = c & ~(KNN_STAGE - 1);
    for (int j = 0; j < (c & (KNN_STAGE - 1)); ++j)
      st[j * 256 + lane] = cand[row_at + c0 + j];
  }
};

template <int F, int G>
AXB_D int knn_collect_walk(const TreeView &tv, float qx, float qy, float qz, float &R,
                           float &R_hi, float R_lo, float shrink, uint2 *__restrict__ cand,
                           uint32_t row_at, int cap, CandStage &stage, uint32_t &v_int,
                           uint32_t &v_leaf)
{
  constexpr uint32_t SL = leaf_slots<G>(), SI = internal_slots<F>();
  float thr = sqrt_threshold(R);
  if constexpr (G == GEOM_POINTS)
  {
    // one record per step, leaf and internal node through the same code
    // (the closest point of a leaf is the point), as in rope_walk_points.
    // A full row does not end the walk: the radius shrinks (by `shrink`,
    // up to 3 times, staying above R_lo, a radius known to hold fewer than
    // k leaves) and the row keeps the entries within the new radius; every
    // leaf passed over so far lay outside the old radius, so the row still
    // holds every leaf seen within the new one, and the walk goes on (R_hi:
    // the last radius known to hold at least the row's capacity).
    // Candidates go through the lane's LDS stage (cap is a multiple of 8, so
    // a full row has been written out whole).
    int c = 0;
    uint32_t p = 0;
    const uint32_t end = tv.total;
    uint32_t step = 1; // 0 after a zero rope: a corrupt node array
    WalkFrame fr;
    walk_frame_init(tv, fr);
    while (true)
    {
      // the walk proper, until its end or a leaf within R finds the row
      // full: c reaches cap + 1 and that leaf waits in stage entry cap (or a
      // zero rope ends it).  The loop carries nothing but p, c, the step and
      // the frame; the rare shrink below stays out of it, so that it carries
      // neither R nor its bounds.
      while (step != 0 && p < end && c <= cap)
      {
        const uint4 s0 = tv.slots[p];
        const uint4 s1 = F == NODES_FP32 ? tv.slots[p + 1] : s0;
        const int32_t w = (int32_t)s0.w;
        const bool leaf = w < 0;
        Box6 bx;
        walk_node_box<F>(s0, s1, fr, bx);
        const float cx = leaf ? bits_f(s0.x) : clamp_box(qx, bx.lo[0], bx.hi[0]);
        const float cy = leaf ? bits_f(s0.y) : clamp_box(qy, bx.lo[1], bx.hi[1]);
        const float cz = leaf ? bits_f(s0.z) : clamp_box(qz, bx.lo[2], bx.hi[2]);
        const float d2 = dist2_pp(qx, qy, qz, cx, cy, cz);
        const bool pass = d2 <= thr;
        v_leaf += leaf;
        v_int += !leaf;
        walk_frame_next<F>(tv, !leaf && pass, bx, fr);
        step = leaf ? 1u : (pass ? SI : (uint32_t)w);
        p += step;
        // A candidate goes to the stage; every lane writes its next free
        // stage entry each step (without a candidate that entry stays
        // free), which keeps the step free of a branch the wavefront would
        // take at almost every step (some lane has a candidate).
        const bool hit = leaf && pass;
        stage.put(c, make_uint2((uint32_t)~w, f_bits(d2)));
        c += hit;
        if (hit && (c & (KNN_STAGE - 1)) == 0)
          stage.flush(cand, row_at, c - KNN_STAGE);
      }
      if (step == 0 || c <= cap)
        break;
      // the row's cap entries are all written out (cap is a multiple of the
      // stage's size); the leaf that found it full is in the stage
      const uint2 held = stage.get(cap);
      c = cap;
      for (int t = 0; t < 3 && c == cap && R * shrink > R_lo; ++t)
      {
        R_hi = R;
        R = R * shrink;
        thr = sqrt_threshold(R);
        int kept = 0;
        for (int j = 0; j < c; ++j)
        {
          const uint2 e = cand[row_at + j];
          if (bits_f(e.y) <= thr)
            cand[row_at + kept++] = e;
        }
        c = kept;
      }
      if (c == cap)
        return cap + 1; // still full: left to the next round
      stage.reload(cand, row_at, c);
      if (bits_f(held.y) <= thr)
      {
        stage.put(c, held);
        ++c;
        if ((c & (KNN_STAGE - 1)) == 0)
          stage.flush(cand, row_at, c - KNN_STAGE);
      }
    }
    if (step == 0)
    {
      raise_error(tv.err, KERR_LAYOUT);
      return c;
    }
    if (c & (KNN_STAGE - 1)) // the last, partial chunk (the rest of it unused)
      stage.flush(cand, row_at, c & ~(KNN_STAGE - 1));
    return c;
  }
  // Box leaves: one record per step, leaf and internal node through the
  // same code, as in rope_walk_boxes (the box is the leaf's exact one or the
  // node's decoded one, picked by selects).  The walk stops at cap + 1
  // candidates (the next round shrinks the radius).
  int c = 0;
  uint32_t p = 0;
  const uint32_t end = tv.total;
  uint32_t step = 1; // 0 after a zero rope (reported below), which ends the loop
  WalkFrame fr;
  walk_frame_init(tv, fr);
  while (step != 0 && p < end && c <= cap)
  {
    const uint4 s0 = tv.slots[p];
    const uint4 s1 = tv.slots[p + 1];
    const int32_t w = (int32_t)s0.w;
    const bool leaf = w < 0;
    Box6 bx;
    walk_node_box<F>(s0, s1, fr, bx);
    if (F == NODES_Q16)
    {
      bx.lo[0] = leaf ? bits_f(s0.x) : bx.lo[0];
      bx.lo[1] = leaf ? bits_f(s0.y) : bx.lo[1];
      bx.lo[2] = leaf ? bits_f(s0.z) : bx.lo[2];
      bx.hi[0] = leaf ? bits_f(s1.x) : bx.hi[0];
      bx.hi[1] = leaf ? bits_f(s1.y) : bx.hi[1];
      bx.hi[2] = leaf ? bits_f(s1.z) : bx.hi[2];
    }
    const float d2 = dist2_pnode(qx, qy, qz, bx); // leaf_d2 for a leaf
    const bool pass = d2 <= thr;
    v_leaf += leaf;
    v_int += !leaf;
    if (leaf && pass)
    {
      if (c < cap)
        cand[row_at + c] = make_uint2((uint32_t)~w, f_bits(d2));
      ++c;
    }
    walk_frame_next<F>(tv, !leaf && pass, bx, fr);
    step = leaf ? SL : (pass ? SI : (uint32_t)w);
    p += step;
  }
  if (step == 0)
    raise_error(tv.err, KERR_LAYOUT);
  return c;
}

// One lane per query: seed radius, then up to KNN_ROUNDS collect walks until
// one finds between k and cap leaves within the radius (then every point
// within the true k-th distance is in the row), followed by the exact top-k
// selection of that row while it is still in L2.  A walk whose row fills up
// shrinks its radius by the volume ratio (2k / (cap + 1))^(1/3) and goes on
// (points; boxes stop at cap + 1 and the next round shrinks); with fewer
// than k leaves the radius was below the k-th distance and the next round
// bisects towards the last radius known to hold enough.  Queries still unresolved go to the exact fallback.  The row
// is addressed as cand + row_at (a scalar base and a 32-bit offset): the
// walk then fits the 64 registers of full occupancy without spilling.
constexpr int KNN_ROUNDS = 6;

template <int F, int G, int K, bool COUNT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((K <= 16 && G == GEOM_POINTS) ? 8 : 1))) void k_knn_collect(
    TreeView tv, const float *__restrict__ pts, int64_t t0, int64_t tn, int k, int W,
    const uint32_t *__restrict__ order, uint2 *__restrict__ cand, int cap, float shrink,
    int max_rounds, uint32_t *__restrict__ ovf, uint32_t *__restrict__ ovf_n,
    int32_t *__restrict__ oi, float *__restrict__ od)
{
  __shared__ __align__(16) uint2 s_stage[G == GEOM_POINTS ? KNN_STAGE * 256 : 1];
  const int64_t g = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (g >= tn)
    return;
  CandStage stage{s_stage, (int)threadIdx.x};
  const int64_t t = t0 + g;
  const int64_t q = order ? (int64_t)order[t] : t;
  const float qx = pts[3 * q], qy = pts[3 * q + 1], qz = pts[3 * q + 2];
  const int m = k < tv.n ? k : tv.n;
  uint32_t v_int = 0, v_leaf = 0;
  float R = knn_seed<F, G, K>(tv, qx, qy, qz, k, W, v_int, v_leaf,
                             G == GEOM_POINTS ? s_stage : nullptr);
  const uint32_t row_at = (uint32_t)(g * cap); // batch * cap + cap <= 2^32 - 1
  int c = 0;
  bool done = false;
  int rounds = 0;
  float R_hi = R;   // radius known to hold >= k points
  float R_lo = 0.f; // radius known to hold < k points
  for (int round = 0; round < max_rounds; ++round)
  {
    ++rounds;
    c = knn_collect_walk<F, G>(tv, qx, qy, qz, R, R_hi, R_lo, shrink, cand, row_at, cap, stage,
                               v_int, v_leaf);
    if (c >= m && c <= cap)
    {
      done = true;
      break;
    }
    if (c < m)
    {
      R_lo = R;
      R = 0.5f * (R_lo + R_hi);
      if (!(R > R_lo && R < R_hi))
        R = R_hi;
    }
    else
    {
      // c == cap + 1: shrink towards 2m leaves, staying above R_lo
      R_hi = R;
      float next = R * shrink;
      R = next > R_lo ? next : 0.5f * (R_lo + R_hi);
    }
  }
  if (done) // exact top-k of the row (every leaf within the k-th distance is in it)
    select_row<K>(cand, row_at, c, m, oi, od, q * m);
  else
    ovf[atomicAdd(ovf_n, 1u)] = (uint32_t)t;
  if (COUNT && tv.visits)
  {
    atomicAdd(tv.visits + 0, 1ull);
    atomicAdd(tv.visits + 1, (unsigned long long)v_int);
    atomicAdd(tv.visits + 2, (unsigned long long)v_leaf);
    atomicAdd(tv.visits + 3, (unsigned long long)c);
    atomicAdd(tv.visits + 8, (unsigned long long)rounds);
    if (rounds > 1)
      atomicAdd(tv.visits + 9, 1ull);
  }
}

template <int F, int G, int K>
__global__ __launch_bounds__(256) void k_knn_fallback(TreeView tv, const float *__restrict__ pts,
                                                      int k, const uint32_t *__restrict__ order,
                                                      const uint32_t *__restrict__ ovf,
                                                      const uint32_t *__restrict__ ovf_n,
                                                      int32_t *__restrict__ oi,
                                                      float *__restrict__ od)
{
  __shared__ float s_d2[KNN_BUF * 256];
  __shared__ int32_t s_id[KNN_BUF * 256];
  const uint32_t cnt = *ovf_n;
  const int m = k < tv.n ? k : tv.n;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt;
       i += gridDim.x * blockDim.x)
  {
    const int64_t t = ovf[i];
    const int64_t q = order ? (int64_t)order[t] : t;
    const float qx = pts[3 * q], qy = pts[3 * q + 1], qz = pts[3 * q + 2];
    CandBuf B{s_d2, s_id, 256, (int)threadIdx.x, 0};
    RegList<K> L;
    L.init(k);
    uint32_t v_int = 0, v_leaf = 0;
    const int W = min(tv.n, max(4 * k, 16));
    knn_traverse<F, G>(tv, qx, qy, qz, m, W, L, B, v_int, v_leaf);
    if (tv.visits)
    {
      atomicAdd(tv.visits + 4, 1ull);
      atomicAdd(tv.visits + 5, (unsigned long long)v_int);
      atomicAdd(tv.visits + 6, (unsigned long long)v_leaf);
      atomicAdd(tv.visits + 7, (unsigned long long)m);
    }
    L.write(oi + q * m, od ? od + q * m : nullptr, m);
  }
}

template <int F, int G>
__global__ __launch_bounds__(256) void k_knn_heap(TreeView tv,
                                                  const float *__restrict__ pts,
                                                  int64_t nq, int k,
                                                  const uint32_t *__restrict__ order,
                                                  int32_t *__restrict__ out_idx,
                                                  float *__restrict__ out_dist)
{
  __shared__ float s_d2[KNN_BUF * 256];
  __shared__ int32_t s_id[KNN_BUF * 256];
  CandBuf B{s_d2, s_id, 256, (int)threadIdx.x, 0};
  const int64_t t = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= nq)
    return;
  const int64_t q = order ? (int64_t)order[t] : t;
  const int m = k < tv.n ? k : tv.n;
  const float qx = pts[3 * q], qy = pts[3 * q + 1], qz = pts[3 * q + 2];
  RowHeap L;
  L.init(m, out_idx + q * m, out_dist + q * m);
  uint32_t v_int = 0, v_leaf = 0;
  knn_traverse<F, G>(tv, qx, qy, qz, m, min(tv.n, 2 * m), L, B, v_int, v_leaf);
  if (tv.visits)
  {
    atomicAdd(tv.visits + 0, 1ull);
    atomicAdd(tv.visits + 1, (unsigned long long)v_int);
    atomicAdd(tv.visits + 2, (unsigned long long)v_leaf);
    atomicAdd(tv.visits + 3, (unsigned long long)m);
  }
  L.finish(m);
}

__global__ void k_knn_offsets(int64_t *offsets, int64_t nq, int64_t per)
{
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= nq)
    offsets[i] = i * per;
}

// Morton-32 code of predicate centroids in the tree's scene box
// (SpaceFillingCurves.hpp:31-43, CrsGraphWrapperImpl.hpp:407-419) plus the
// digit histograms of the 4 radix passes.
template <int PK>
__global__ __launch_bounds__(256) void k_pred_morton32(
    const float *__restrict__ preds, int64_t nq, const float *__restrict__ scene, int shift,
    uint32_t *__restrict__ keys, uint32_t *__restrict__ hist)
{
  __shared__ uint32_t s_h[4 * 256];
  for (int j = threadIdx.x; j < 4 * 256; j += blockDim.x)
    s_h[j] = 0;
  const float a0 = scene[0], a1 = scene[1], a2 = scene[2];
  const float b0 = scene[3], b1 = scene[4], b2 = scene[5];
  __syncthreads();
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += stride)
  {
    float cx, cy, cz;
    if (PK == PRED_RAYS) // a ray's centroid is its origin (ArborX_Ray.hpp:76-80)
    {
      const float *v = preds + 6 * i;
      cx = v[0], cy = v[1], cz = v[2];
    }
    else if (PK == PRED_BOXES)
    {
      const float *v = preds + 6 * i;
      cx = (v[0] + v[3]) / 2, cy = (v[1] + v[4]) / 2, cz = (v[2] + v[5]) / 2;
    }
    else
    {
      const float *v = preds + (PK == PRED_SPHERES ? 4 : 3) * i;
      cx = v[0], cy = v[1], cz = v[2];
    }
    const uint32_t k = morton32(translate_and_scale(cx, a0, b0),
                                translate_and_scale(cy, a1, b1),
                                translate_and_scale(cz, a2, b2)) >>
                       shift;
    keys[i] = k;
#pragma unroll
    for (int p = 0; p < 4; ++p)
      atomicAdd(&s_h[p * 256 + ((k >> (8 * p)) & 255u)], 1u);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < 4 * 256; j += blockDim.x)
    if (s_h[j])
      atomicAdd(&hist[j], s_h[j]);
}

// Buffered first pass -> CRS: row q's results sit at buf + q * buffer when
// they fit (CrsGraphWrapperImpl.hpp:296-318, copy_valid_values); rows that
// did not fit are written by a second traversal of those queries.
// Scratch rows that fit -> their CRS rows.  One wavefront per run of
// COMPACT_ROWS consecutive rows, whose CRS rows are one contiguous range:
// it writes that range in order, lane l of each step at position
// start + l + 64 i (whole lines, no partial writes at the rows' unaligned
// starts), each lane following the row its position falls in through the
// run's offsets in LDS.  Positions of rows that did not fit are passed over
// (the second traversal writes them).
constexpr int COMPACT_ROWS = 64;
__global__ __launch_bounds__(256) void k_compact(const int32_t *__restrict__ buf, int64_t nq,
                                                 int32_t buffer,
                                                 const int64_t *__restrict__ offsets,
                                                 int32_t *__restrict__ out)
{
  __shared__ int64_t s_off[4][COMPACT_ROWS + 1];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t q0 = ((int64_t)blockIdx.x * 4 + wave) * COMPACT_ROWS;
  const int rows = q0 >= nq ? 0 : (int)(nq - q0 < COMPACT_ROWS ? nq - q0 : COMPACT_ROWS);
  int64_t *so = s_off[wave];
  // so[r] = offsets[q0 + r] for r <= rows, then offsets[q0 + rows] (empty
  // rows); offsets holds nq + 1 entries
  if (rows > 0)
  {
    so[lane] = offsets[q0 + (lane < rows ? lane : rows)];
    if (lane == 0)
      so[COMPACT_ROWS] = offsets[q0 + rows];
  }
  __syncthreads();
  if (rows == 0)
    return;
  const int64_t end = so[rows];
  int r = 0;
  for (int64_t j = so[0] + lane; j < end;)
  {
    // j < end = so[rows] and so is non-decreasing: r stops below rows
    while (so[r + 1] <= j)
      ++r;
    const int64_t start = so[r], stop = so[r + 1];
    if (stop - start <= buffer)
    {
      out[j] = buf[(q0 + r) * (int64_t)buffer + (j - start)];
      j += 64;
    }
    else // past the row, keeping j's place in the line pattern
      j += (stop - j + 63) / 64 * 64;
  }
}

// Rows whose result count exceeds the buffer, listed in processing
// (Morton) order: one atomic per wavefront, lanes in order behind it.  The
// list holds the rows, or the processing indices when emit_index is set.
// over_flags: one bit per row (set by the offsets' scan), 1/64 of the
// offsets' bytes, so that the lookups in processing order stay in L2.
__global__ void k_overflow_list(const uint8_t *__restrict__ over_flags,
                                const uint32_t *__restrict__ order, int64_t nq, int emit_index,
                                uint32_t *__restrict__ ovf, uint32_t *__restrict__ ovf_n)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  uint32_t q = 0;
  bool f = false;
  if (t < nq)
  {
    q = order ? order[t] : (uint32_t)t;
    f = (over_flags[q >> 3] >> (q & 7)) & 1;
  }
  const uint64_t m = __ballot(f);
  if (m == 0)
    return;
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader)
    base = atomicAdd(ovf_n, (uint32_t)__popcll(m));
  base = __shfl(base, leader, 64);
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  if (f)
    ovf[base + __popcll(m & lt)] = emit_index ? (uint32_t)t : q;
}

// ------------------------------------------------------------ half -> full
// ExpandHalfToFull.hpp:23-68.  Counts: row i gains its own entries and one
// for each row that lists it.  Entries outside [0, n) raise KERR_LAYOUT.
__global__ void k_h2f_count(const int64_t *__restrict__ offsets,
                            const int32_t *__restrict__ indices, int64_t n,
                            int64_t *__restrict__ cnt, uint32_t *err)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const int64_t b = offsets[i], e = offsets[i + 1];
  if (e > b)
    atomicAdd((unsigned long long *)(cnt + i), (unsigned long long)(e - b));
  for (int64_t p = b; p < e; ++p)
  {
    const int32_t k = indices[p];
    if (k < 0 || k >= n)
    {
      raise_error(err, KERR_LAYOUT);
      continue;
    }
    atomicAdd((unsigned long long *)(cnt + k), 1ull);
  }
}

__global__ void k_h2f_cursor(const int64_t *__restrict__ offsets,
                             const int64_t *__restrict__ out_offsets, int64_t n,
                             int64_t *__restrict__ cursor)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    cursor[i] = out_offsets[i] + (offsets[i + 1] - offsets[i]);
}

// Row i of the full list: its own entries first (in order), then the
// mirrored ones, appended through the row's cursor.
__global__ void k_h2f_fill(const int64_t *__restrict__ offsets,
                           const int32_t *__restrict__ indices, int64_t n,
                           const int64_t *__restrict__ out_offsets, int64_t *__restrict__ cursor,
                           int32_t *__restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const int64_t b = offsets[i], e = offsets[i + 1], o = out_offsets[i];
  for (int64_t p = b; p < e; ++p)
  {
    const int32_t k = indices[p];
    if (k < 0 || k >= n)
      continue;
    out[o + (p - b)] = k;
    const int64_t at = (int64_t)atomicAdd((unsigned long long *)(cursor + k), 1ull);
    if (at < out_offsets[k + 1])
      out[at] = (int32_t)i;
  }
}

} // namespace

// ------------------------------------------------------------ host launchers
inline unsigned grid_of(int64_t n) { return (unsigned)div_up(n, 256); }


template <int F, int G, int PK, bool COUNT>
static hipError_t launch_spatial_fgc(const TreeView &tv, const float *preds, int64_t nq,
                                     float radius, const uint32_t *order, int64_t *offsets,
                                     int32_t *out, int32_t buffer, int mode, hipStream_t s)
{
  if (mode == MODE_COUNT)
    AXB_LAUNCH((k_spatial<F, G, PK, MODE_COUNT, false, COUNT>), dim3(grid_of(nq)), dim3(256), 0,
               s, tv, preds, nq, radius, order, offsets, out, buffer);
  else if (mode == MODE_FILL)
    AXB_LAUNCH((k_spatial<F, G, PK, MODE_FILL, false, COUNT>), dim3(grid_of(nq)), dim3(256), 0,
               s, tv, preds, nq, radius, order, offsets, out, buffer);
  else
    AXB_LAUNCH((k_spatial<F, G, PK, MODE_BUFFER, false, COUNT>), dim3(grid_of(nq)), dim3(256),
               0, s, tv, preds, nq, radius, order, offsets, out, buffer);
  return hipGetLastError();
}
template <int F, int G, int PK>
static hipError_t launch_spatial_fg(const TreeView &tv, const float *preds, int64_t nq,
                                    float radius, const uint32_t *order, int64_t *offsets,
                                    int32_t *out, int32_t buffer, int mode, hipStream_t s)
{
  return tv.visits ? launch_spatial_fgc<F, G, PK, true>(tv, preds, nq, radius, order, offsets,
                                                         out, buffer, mode, s)
                   : launch_spatial_fgc<F, G, PK, false>(tv, preds, nq, radius, order, offsets,
                                                          out, buffer, mode, s);
}

template <int PK>
static hipError_t launch_spatial_p(const TreeView &tv, int F, int G, const float *preds,
                                   int64_t nq, float radius, const uint32_t *order,
                                   int64_t *offsets, int32_t *out, int32_t buffer, int mode,
                                   hipStream_t s)
{
  if (F == NODES_FP32 && G == GEOM_POINTS)
    return launch_spatial_fg<NODES_FP32, GEOM_POINTS, PK>(tv, preds, nq, radius, order, offsets,
                                                          out, buffer, mode, s);
  if (F == NODES_FP32 && G == GEOM_BOXES)
    return launch_spatial_fg<NODES_FP32, GEOM_BOXES, PK>(tv, preds, nq, radius, order, offsets,
                                                         out, buffer, mode, s);
  if (F == NODES_Q16 && G == GEOM_POINTS)
    return launch_spatial_fg<NODES_Q16, GEOM_POINTS, PK>(tv, preds, nq, radius, order, offsets,
                                                         out, buffer, mode, s);
  return launch_spatial_fg<NODES_Q16, GEOM_BOXES, PK>(tv, preds, nq, radius, order, offsets,
                                                      out, buffer, mode, s);
}

hipError_t launch_spatial(const TreeView &tv, int node_format, int geometry, int pred_kind,
                          const float *preds, int64_t nq, float radius, const uint32_t *order,
                          int64_t *offsets, int32_t *out, int32_t buffer, int mode,
                          hipStream_t s)
{
  if (nq <= 0)
    return hipSuccess;
  if (pred_kind == PRED_RAYS)
  {
    // box trees only (the caller checks)
    if (node_format == NODES_FP32)
      return launch_spatial_fg<NODES_FP32, GEOM_BOXES, PRED_RAYS>(tv, preds, nq, radius, order,
                                                                  offsets, out, buffer, mode, s);
    return launch_spatial_fg<NODES_Q16, GEOM_BOXES, PRED_RAYS>(tv, preds, nq, radius, order,
                                                               offsets, out, buffer, mode, s);
  }
  if (pred_kind == PRED_SPHERES)
    return launch_spatial_p<PRED_SPHERES>(tv, node_format, geometry, preds, nq, radius, order,
                                          offsets, out, buffer, mode, s);
  if (pred_kind == PRED_BOXES)
    return launch_spatial_p<PRED_BOXES>(tv, node_format, geometry, preds, nq, radius, order,
                                        offsets, out, buffer, mode, s);
  return launch_spatial_p<PRED_RADIUS>(tv, node_format, geometry, preds, nq, radius, order,
                                       offsets, out, buffer, mode, s);
}

// Shape of the collect launches for k <= 64 (the candidate-row capacity and
// the radius factor after an overflowing round).
struct KnnShape
{
  int W, cap;
  float shrink;
};
static KnnShape knn_shape(int k, int n)
{
  KnnShape sh;
  // seed window: 6k leaves (measured at 10M points, k = 16, Q16 nodes:
  // 3k 11.2 ms uniform / 15.5 ms Gaussian mixture, 4k 10.9 / 14.5, 6k
  // 11.0 / 13.8 -- a tighter seed radius against a longer seed scan)
  sh.W = std::min(n, std::max(6 * k, 16));
  // row capacity: about 6k, a multiple of 8 (10M points, k = 16: 4k 13.1 ms
  // uniform / 20.0 ms Gaussian mixture, 6k 12.85 / 20.9, 8k 13.1 / 23.2 --
  // a full row costs an in-place shrink, a long row a longer selection)
  sh.cap = 6 * k > 64 ? (6 * k + 7) / 8 * 8 : 64;
  // radius factor after an overflowing round (cap + 1 leaves within R):
  // aim at 2m leaves, m = min(k, n)
  sh.shrink = std::cbrt((float)(2 * std::min(k, n)) / (float)(sh.cap + 1));
  return sh;
}

int knn_row_capacity(int k, int n) { return knn_shape(k, n).cap; }

template <int F, int G, int K>
static hipError_t knn_run(const TreeView &tv, const float *pts, int64_t nq, int k,
                          const uint32_t *order, int32_t *oi, float *od, const KnnScratch &ws,
                          hipStream_t s)
{
  const KnnShape sh = knn_shape(k, tv.n);
  const int lim = debug_limits().knn_rounds;
  const int rounds = lim > 0 ? lim : (lim < 0 ? 0 : KNN_ROUNDS);
  hipError_t e = fill_u32(ws.ovf_n, 0u, 1, s);
  for (int64_t t0 = 0; e == hipSuccess && t0 < nq; t0 += ws.batch)
  {
    const int64_t tn = std::min(ws.batch, nq - t0);
    if (tv.visits)
      AXB_LAUNCH((k_knn_collect<F, G, K, true>), dim3(grid_of(tn)), dim3(256), 0, s, tv, pts, t0,
                 tn, k, sh.W, order, ws.cand, sh.cap, sh.shrink, rounds, ws.ovf, ws.ovf_n, oi, od);
    else
      AXB_LAUNCH((k_knn_collect<F, G, K, false>), dim3(grid_of(tn)), dim3(256), 0, s, tv, pts,
                 t0, tn, k, sh.W, order, ws.cand, sh.cap, sh.shrink, rounds, ws.ovf, ws.ovf_n, oi,
                 od);
    e = hipGetLastError();
  }
  // unresolved queries of every batch, in one launch at the end: each of
  // them is one lane walking alone, so launches of a few lanes per batch
  // would each cost the longest such walk
  if (e == hipSuccess)
  {
    AXB_LAUNCH((k_knn_fallback<F, G, K>), dim3((unsigned)std::min<int64_t>(grid_of(nq), 512)),
               dim3(256), 0, s, tv, pts, k, order, ws.ovf, ws.ovf_n, oi, od);
    e = hipGetLastError();
  }
  return e;
}

template <int F, int G, int K>
static hipError_t knn_pipeline(const TreeView &tv, const float *pts, int64_t nq, int k,
                               const uint32_t *order, int32_t *oi, float *od, hipStream_t s)
{
  const int cap = knn_shape(k, tv.n).cap;
  // queries per launch: candidate rows of up to 8 GiB, row offsets within
  // 32 bits (10M points, k = 16: 4 launches 13.3 ms uniform / 20.9 ms
  // Gaussian mixture, 2 launches 12.7 / 17.5, 1 launch 12.4 / 14.8 -- every
  // launch ends with a tail of the slowest wavefronts)
  int64_t batch = std::max<int64_t>(
      1, std::min<int64_t>(nq, std::min<int64_t>((int64_t(1) << 33) / (cap * 8),
                                                 (int64_t(UINT32_MAX) - cap) / cap)));
  const int64_t lim = debug_limits().knn_batch_max;
  if (lim > 0)
    batch = std::min(batch, lim);
  // smaller launches when the device cannot spare that much
  WsLease rows;
  while (true)
  {
    rows.acquire(s, WS_KNN_CAND, (size_t)batch * cap * sizeof(uint2));
    if (rows.p || batch <= 65536)
      break;
    batch = (batch + 1) / 2;
  }
  KnnScratch ws{(uint2 *)rows.p, batch, (uint32_t *)dalloc((size_t)nq * sizeof(uint32_t), s),
                (uint32_t *)dalloc(sizeof(uint32_t), s)};
  hipError_t e = (ws.cand && ws.ovf && ws.ovf_n) ? hipSuccess : hipErrorOutOfMemory;
  if (e == hipSuccess)
    e = knn_run<F, G, K>(tv, pts, nq, k, order, oi, od, ws, s);
  dfree(ws.ovf, s);
  dfree(ws.ovf_n, s);
  return e;
}

template <int F, int G>
static hipError_t launch_knn_fg(const TreeView &tv, const float *pts, int64_t nq, int k,
                                const uint32_t *order, int32_t *oi, float *od, hipStream_t s)
{
  if (k <= 1)
    return knn_pipeline<F, G, 1>(tv, pts, nq, k, order, oi, od, s);
  if (k <= 2)
    return knn_pipeline<F, G, 2>(tv, pts, nq, k, order, oi, od, s);
  if (k <= 4)
    return knn_pipeline<F, G, 4>(tv, pts, nq, k, order, oi, od, s);
  if (k <= 8)
    return knn_pipeline<F, G, 8>(tv, pts, nq, k, order, oi, od, s);
  if (k <= 16)
    return knn_pipeline<F, G, 16>(tv, pts, nq, k, order, oi, od, s);
  if (k <= 32)
    return knn_pipeline<F, G, 32>(tv, pts, nq, k, order, oi, od, s);
  if (k <= 64)
    return knn_pipeline<F, G, 64>(tv, pts, nq, k, order, oi, od, s);
  AXB_LAUNCH((k_knn_heap<F, G>), dim3(grid_of(nq)), dim3(256), 0, s, tv, pts, nq, k, order, oi,
             od);
  return hipGetLastError();
}

// od must be non-null when k > 64 (the heap keeps its distances there).
hipError_t launch_knn(const TreeView &tv, int node_format, int geometry, const float *pts,
                      int64_t nq, int k, const uint32_t *order, int32_t *oi, float *od,
                      hipStream_t s)
{
  if (nq <= 0 || k < 1 || tv.n == 0)
    return hipSuccess;
  if (node_format == NODES_FP32 && geometry == GEOM_POINTS)
    return launch_knn_fg<NODES_FP32, GEOM_POINTS>(tv, pts, nq, k, order, oi, od, s);
  if (node_format == NODES_FP32 && geometry == GEOM_BOXES)
    return launch_knn_fg<NODES_FP32, GEOM_BOXES>(tv, pts, nq, k, order, oi, od, s);
  if (node_format == NODES_Q16 && geometry == GEOM_POINTS)
    return launch_knn_fg<NODES_Q16, GEOM_POINTS>(tv, pts, nq, k, order, oi, od, s);
  return launch_knn_fg<NODES_Q16, GEOM_BOXES>(tv, pts, nq, k, order, oi, od, s);
}

template <int F>
static hipError_t launch_ray_slots_f(const TreeView &tv, const float *rays, int64_t nq,
                                     const uint32_t *order, int64_t *offsets, int32_t *out,
                                     int32_t buffer, int mode, hipStream_t s)
{
  constexpr int G = GEOM_BOXES, PK = PRED_RAYS;
  if (mode == MODE_COUNT)
    AXB_LAUNCH((k_spatial<F, G, PK, MODE_COUNT, true>), dim3(grid_of(nq)), dim3(256), 0, s, tv,
               rays, nq, 0.f, order, offsets, out, buffer);
  else if (mode == MODE_FILL)
    AXB_LAUNCH((k_spatial<F, G, PK, MODE_FILL, true>), dim3(grid_of(nq)), dim3(256), 0, s, tv,
               rays, nq, 0.f, order, offsets, out, buffer);
  else
    AXB_LAUNCH((k_spatial<F, G, PK, MODE_BUFFER, true>), dim3(grid_of(nq)), dim3(256), 0, s, tv,
               rays, nq, 0.f, order, offsets, out, buffer);
  return hipGetLastError();
}

hipError_t launch_ray_slots(const TreeView &tv, int node_format, const float *rays, int64_t nq,
                            const uint32_t *order, int64_t *offsets, int32_t *out,
                            int32_t buffer, int mode, hipStream_t s)
{
  if (nq <= 0)
    return hipSuccess;
  if (node_format == NODES_FP32)
    return launch_ray_slots_f<NODES_FP32>(tv, rays, nq, order, offsets, out, buffer, mode, s);
  return launch_ray_slots_f<NODES_Q16>(tv, rays, nq, order, offsets, out, buffer, mode, s);
}

hipError_t launch_row_prefix_counts(const int64_t *offsets, int64_t nq, int32_t k,
                                    int64_t *counts, hipStream_t s)
{
  AXB_LAUNCH(k_row_prefix_counts, dim3(grid_of(nq + 1)), dim3(256), 0, s, offsets, nq, k,
             counts);
  return hipGetLastError();
}

hipError_t launch_row_prefix_copy(const int64_t *offsets, const int32_t *rows, int64_t nq,
                                  const int64_t *out_offsets, int32_t *out, hipStream_t s)
{
  if (nq <= 0)
    return hipSuccess;
  AXB_LAUNCH(k_row_prefix_copy, dim3(grid_of(nq)), dim3(256), 0, s, offsets, rows, nq,
             out_offsets, out);
  return hipGetLastError();
}

hipError_t launch_ray_sort_rows(const TreeView &tv, const float *rays, int64_t nq,
                                const int64_t *offsets, int32_t *rows, hipStream_t s)
{
  if (nq <= 0)
    return hipSuccess;
  AXB_LAUNCH(k_ray_sort_rows<0>, dim3(grid_of(nq)), dim3(256), 0, s, tv, rays, nq, offsets,
             rows);
  AXB_LAUNCH(k_ray_sort_rows<1>, dim3(grid_of(nq)), dim3(256), 0, s, tv, rays, nq, offsets,
             rows);
  AXB_LAUNCH(k_ray_sort_rows<2>, dim3(grid_of(nq)), dim3(256), 0, s, tv, rays, nq, offsets,
             rows);
  return hipGetLastError();
}

template <int F, int G>
static hipError_t launch_ordered_rows_fg(const TreeView &tv, int pred_kind, const float *preds,
                                         int64_t nq, int32_t *rows, hipStream_t s)
{
  AXB_LAUNCH((k_leaf_slot_rows<F, G>), dim3(grid_of(nq * (int64_t)tv.n)), dim3(256), 0, s, tv,
             nq, rows);
  if (pred_kind == PRED_SPHERES)
    AXB_LAUNCH((k_ordered_sort_rows<G, PRED_SPHERES>), dim3(grid_of(nq)), dim3(256), 0, s, tv,
               preds, nq, rows);
  else if (pred_kind == PRED_BOXES)
    AXB_LAUNCH((k_ordered_sort_rows<G, PRED_BOXES>), dim3(grid_of(nq)), dim3(256), 0, s, tv,
               preds, nq, rows);
  else
    AXB_LAUNCH((k_ordered_sort_rows<G, PRED_POINTS>), dim3(grid_of(nq)), dim3(256), 0, s, tv,
               preds, nq, rows);
  return hipGetLastError();
}

hipError_t launch_ordered_rows(const TreeView &tv, int node_format, int geometry, int pred_kind,
                               const float *preds, int64_t nq, int32_t *rows, hipStream_t s)
{
  if (nq <= 0 || tv.n <= 0)
    return hipSuccess;
  if (node_format == NODES_Q16)
    return geometry == GEOM_POINTS
               ? launch_ordered_rows_fg<NODES_Q16, GEOM_POINTS>(tv, pred_kind, preds, nq, rows, s)
               : launch_ordered_rows_fg<NODES_Q16, GEOM_BOXES>(tv, pred_kind, preds, nq, rows, s);
  return geometry == GEOM_POINTS
             ? launch_ordered_rows_fg<NODES_FP32, GEOM_POINTS>(tv, pred_kind, preds, nq, rows, s)
             : launch_ordered_rows_fg<NODES_FP32, GEOM_BOXES>(tv, pred_kind, preds, nq, rows, s);
}

template <int F>
static hipError_t launch_ordered_f(const TreeView &tv, const float *rays, int64_t nq,
                                   int32_t limit, const uint32_t *order, int64_t *offsets,
                                   int32_t *out, int32_t buffer, int mode, hipStream_t s)
{
  if (mode == MODE_COUNT)
    AXB_LAUNCH((k_ordered<F, MODE_COUNT>), dim3(grid_of(nq)), dim3(256), 0, s, tv, rays, nq,
               limit, order, offsets, out, buffer);
  else if (mode == MODE_FILL)
    AXB_LAUNCH((k_ordered<F, MODE_FILL>), dim3(grid_of(nq)), dim3(256), 0, s, tv, rays, nq,
               limit, order, offsets, out, buffer);
  else
    AXB_LAUNCH((k_ordered<F, MODE_BUFFER>), dim3(grid_of(nq)), dim3(256), 0, s, tv, rays, nq,
               limit, order, offsets, out, buffer);
  return hipGetLastError();
}

hipError_t launch_ordered(const TreeView &tv, int node_format, const float *rays, int64_t nq,
                          int32_t limit, const uint32_t *order, int64_t *offsets, int32_t *out,
                          int32_t buffer, int mode, hipStream_t s)
{
  if (nq <= 0)
    return hipSuccess;
  if (node_format == NODES_FP32)
    return launch_ordered_f<NODES_FP32>(tv, rays, nq, limit, order, offsets, out, buffer, mode,
                                        s);
  return launch_ordered_f<NODES_Q16>(tv, rays, nq, limit, order, offsets, out, buffer, mode, s);
}

hipError_t launch_make_rays(const float *origins, const float *directions, int64_t n,
                            float *rays, uint32_t *err, hipStream_t s)
{
  if (n <= 0)
    return hipSuccess;
  AXB_LAUNCH(k_make_rays, dim3(grid_of(n)), dim3(256), 0, s, origins, directions, n, rays, err);
  return hipGetLastError();
}

hipError_t launch_knn_offsets(int64_t *offsets, int64_t nq, int64_t per, hipStream_t s)
{
  AXB_LAUNCH(k_knn_offsets, dim3(grid_of(nq + 1)), dim3(256), 0, s, offsets, nq, per);
  return hipGetLastError();
}

hipError_t launch_pred_morton32(int pred_kind, const float *preds, int64_t nq,
                                const float *scene, int shift, uint32_t *keys, uint32_t *hist,
                                hipStream_t s)
{
  const unsigned g = (unsigned)std::min<int64_t>(2048, div_up(nq, 256));
  if (pred_kind == PRED_RAYS)
    AXB_LAUNCH(k_pred_morton32<PRED_RAYS>, dim3(g), dim3(256), 0, s, preds, nq, scene, shift, keys,
               hist);
  else if (pred_kind == PRED_SPHERES)
    AXB_LAUNCH(k_pred_morton32<PRED_SPHERES>, dim3(g), dim3(256), 0, s, preds, nq,
                       scene, shift, keys, hist);
  else if (pred_kind == PRED_BOXES)
    AXB_LAUNCH(k_pred_morton32<PRED_BOXES>, dim3(g), dim3(256), 0, s, preds, nq, scene,
               shift, keys, hist);
  else
    AXB_LAUNCH(k_pred_morton32<PRED_RADIUS>, dim3(g), dim3(256), 0, s, preds, nq,
                       scene, shift, keys, hist);
  return hipGetLastError();
}

hipError_t launch_compact(const int32_t *buf, int64_t nq, int32_t buffer,
                          const int64_t *offsets, int32_t *out, hipStream_t s)
{
  if (nq <= 0)
    return hipSuccess;
  AXB_LAUNCH(k_compact, dim3((unsigned)div_up(nq, 4 * COMPACT_ROWS)), dim3(256), 0, s, buf, nq,
             buffer, offsets, out);
  return hipGetLastError();
}

hipError_t launch_overflow_list(const uint8_t *over_flags, const uint32_t *order, int64_t nq,
                                bool emit_index, uint32_t *ovf, uint32_t *ovf_n, hipStream_t s)
{
  if (nq <= 0)
    return hipSuccess;
  AXB_LAUNCH(k_overflow_list, dim3(grid_of(nq)), dim3(256), 0, s, over_flags, order, nq,
             emit_index ? 1 : 0, ovf, ovf_n);
  return hipGetLastError();
}

template <int F, bool FULL>
static hipError_t launch_nlist_f(const TreeView &tv, float radius, int64_t nl,
                                 const uint32_t *list, int64_t *offsets, int32_t *out,
                                 int32_t buffer, uint32_t *rowof, int mode, hipStream_t s)
{
  if (mode == MODE_COUNT)
    AXB_LAUNCH((k_nlist<F, FULL, MODE_COUNT>), dim3(grid_of(nl)), dim3(256), 0, s, tv, radius,
               nl, list, offsets, out, buffer, rowof);
  else if (mode == MODE_FILL)
    AXB_LAUNCH((k_nlist<F, FULL, MODE_FILL>), dim3(grid_of(nl)), dim3(256), 0, s, tv, radius,
               nl, list, offsets, out, buffer, rowof);
  else
    AXB_LAUNCH((k_nlist<F, FULL, MODE_BUFFER>), dim3(grid_of(nl)), dim3(256), 0, s, tv, radius,
               nl, list, offsets, out, buffer, rowof);
  return hipGetLastError();
}

hipError_t launch_nlist(const TreeView &tv, int node_format, bool full, float radius, int64_t nl,
                        const uint32_t *list, int64_t *offsets, int32_t *out, int32_t buffer,
                        uint32_t *rowof, int mode, hipStream_t s)
{
  if (nl <= 0)
    return hipSuccess;
  if (node_format == NODES_Q16)
    return full ? launch_nlist_f<NODES_Q16, true>(tv, radius, nl, list, offsets, out, buffer,
                                                  rowof, mode, s)
                : launch_nlist_f<NODES_Q16, false>(tv, radius, nl, list, offsets, out, buffer,
                                                   rowof, mode, s);
  return full ? launch_nlist_f<NODES_FP32, true>(tv, radius, nl, list, offsets, out, buffer,
                                                 rowof, mode, s)
              : launch_nlist_f<NODES_FP32, false>(tv, radius, nl, list, offsets, out, buffer,
                                                  rowof, mode, s);
}

hipError_t launch_h2f_count(const int64_t *offsets, const int32_t *indices, int64_t n,
                            int64_t *cnt, uint32_t *err, hipStream_t s)
{
  if (n <= 0)
    return hipSuccess;
  AXB_LAUNCH(k_h2f_count, dim3(grid_of(n)), dim3(256), 0, s, offsets, indices, n, cnt, err);
  return hipGetLastError();
}

hipError_t launch_h2f_fill(const int64_t *offsets, const int32_t *indices, int64_t n,
                           const int64_t *out_offsets, int64_t *cursor, int32_t *out,
                           hipStream_t s)
{
  if (n <= 0)
    return hipSuccess;
  AXB_LAUNCH(k_h2f_cursor, dim3(grid_of(n)), dim3(256), 0, s, offsets, out_offsets, n, cursor);
  AXB_LAUNCH(k_h2f_fill, dim3(grid_of(n)), dim3(256), 0, s, offsets, indices, n, out_offsets,
             cursor, out);
  return hipGetLastError();
}

} // namespace axb
