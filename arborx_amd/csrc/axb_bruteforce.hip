// axb_bruteforce.hip -- the brute-force index on gfx950: every predicate
// against every value (ArborX::BruteForce, src/spatial/ArborX_BruteForce.hpp,
// detail/ArborX_BruteForceImpl.hpp:53-229).
//
// One lane per predicate; all lanes of a wavefront walk the values in the
// same order, so each value is one wavefront-uniform 12- or 24-byte load
// that every lane tests.  Spatial rows come out in value order through the
// RowWriter modes of the tree queries (count, fill, buffered row), so the
// CRS assembly is the tree query's.  Nearest rows use the register
// selection of the tree's kNN over the keys (d2 bits, value index) of all
// values (k <= 64), or a binary heap in the output row above.
#include "axb_internal.hpp"

namespace axb
{

namespace
{

template <int G>
AXB_D void value_records(const float *__restrict__ values, int64_t j, uint4 &s0, uint4 &s1)
{
  const float *v = values + j * (G == GEOM_POINTS ? 3 : 6);
  s0 = make_uint4(f_bits(v[0]), f_bits(v[1]), f_bits(v[2]), 0u);
  s1 = G == GEOM_POINTS ? s0 : make_uint4(f_bits(v[3]), f_bits(v[4]), f_bits(v[5]), 0u);
}

// Predicate `list ? list[t] : t` against values 0 .. n-1, in that order
// (the reference tests every (predicate, value) pair, BruteForceImpl.hpp:
// 127-151; the order inside a row is its callback order, unspecified).
template <int G, int PK, int MODE>
__global__ __launch_bounds__(256) void k_bf_spatial(const float *__restrict__ values, int64_t n,
                                                    const float *__restrict__ preds, int64_t nq,
                                                    float radius,
                                                    const uint32_t *__restrict__ list,
                                                    int64_t *__restrict__ offsets,
                                                    int32_t *__restrict__ out, int32_t buffer,
                                                    uint32_t *__restrict__ err)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nq)
    return;
  const int64_t q = list ? (int64_t)list[t] : t;
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
  RowWriter<MODE> rw;
  rw.init(out, offsets, q, buffer);
  for (int64_t j = 0; j < n; ++j)
  {
    uint4 s0, s1;
    value_records<G>(values, j, s0, s1);
    if (leaf_hit<G, PK>(pr, s0, s1))
      rw.hit((int32_t)j);
  }
  TreeView tv{};
  tv.err = err;
  row_epilogue(tv, rw, offsets, q, 0u, (uint32_t)n);
}

// Keys (d2 bits, value index) of the values for one query point.
template <int G>
struct ValueKeys
{
  const float *__restrict__ values;
  float qx, qy, qz;
  AXB_D uint64_t at(int j) const
  {
    uint4 s0, s1;
    value_records<G>(values, j, s0, s1);
    const float x = bits_f(s0.x), y = bits_f(s0.y), z = bits_f(s0.z);
    float d2;
    if (G == GEOM_POINTS)
      d2 = dist2_pp(qx, qy, qz, x, y, z);
    else
      d2 = dist2_pbox(qx, qy, qz, Box6{{x, y, z}, {bits_f(s1.x), bits_f(s1.y), bits_f(s1.z)}});
    return ((uint64_t)f_bits(d2) << 32) | (uint32_t)j;
  }
  template <int N>
  AXB_D void load(int b, int c, uint64_t (&a)[N]) const
  {
#pragma unroll
    for (int j = 0; j < N; ++j)
      a[j] = b + j < c ? at(b + j) : ~0ull;
  }
};

// k nearest values of each query point, rows of m = min(k, n) in
// (distance, index) order, distance = RN(sqrt(d2)) (Distance.hpp:52-80).
template <int G, int K>
__global__ __launch_bounds__(256) void k_bf_nearest(const float *__restrict__ values, int n,
                                                    const float *__restrict__ pts, int64_t nq,
                                                    int k, int32_t *__restrict__ oi,
                                                    float *__restrict__ od)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq)
    return;
  const int m = k < n ? k : n;
  const ValueKeys<G> src{values, pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]};
  select_keys<K>(src, n, m, oi + q * m, od ? od + q * m : nullptr);
}

// k > 64: binary max-heap on (distance, index) in the output row, sorted
// at the end (the reference's scheme, BruteForceImpl.hpp:190-227).
template <int G>
__global__ __launch_bounds__(256) void k_bf_heap(const float *__restrict__ values, int n,
                                                 const float *__restrict__ pts, int64_t nq, int k,
                                                 int32_t *__restrict__ oi, float *__restrict__ od)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq)
    return;
  const int m = k < n ? k : n;
  const ValueKeys<G> src{values, pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]};
  RowHeap L;
  L.init(m, oi + q * m, od + q * m);
  for (int j = 0; j < n; ++j)
  {
    const uint64_t e = src.at(j);
    const float dist = sqrt_rn(bits_f((uint32_t)(e >> 32)));
    if (L.accepts(dist, j))
      L.insert(dist, j);
  }
  L.finish(m);
}

// Nearest<Ray> over boxes (BruteForceImpl.hpp:155-229 with distance(Ray,
// Box), ArborX_Ray.hpp:429-444): row q = the first m = min(k, hits) boxes
// of ray q by (distance, index); boxes the ray misses are at +inf and never
// qualify.  A first sweep counts the hits; the second keeps a max-heap of m
// value indices in the row itself (keys recomputed from the values), then
// sorts it.  MODE_BUFFER keeps the row in the scratch row when m fits.
struct RayKeys
{
  const float *__restrict__ values;
  RayPred ray;
  AXB_D uint64_t at(int32_t j) const
  {
    const float *v = values + 6 * (int64_t)j;
    const Box6 b = {{v[0], v[1], v[2]}, {v[3], v[4], v[5]}};
    return ordered_key(ray_distance(ray, b), (uint32_t)j);
  }
};

AXB_D void heap_sift_down(const RayKeys &src, int32_t *h, int j, int size, int32_t v,
                          uint64_t kv)
{
  while (true)
  {
    int c = 2 * j + 1;
    if (c >= size)
      break;
    uint64_t kc = src.at(h[c]);
    if (c + 1 < size)
    {
      const uint64_t kr = src.at(h[c + 1]);
      if (kr > kc)
        ++c, kc = kr;
    }
    if (kv >= kc)
      break;
    h[j] = h[c];
    j = c;
  }
  h[j] = v;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_bf_ray_nearest(const float *__restrict__ values, int n,
                                                        const float *__restrict__ rays,
                                                        int64_t nq, int k,
                                                        const uint32_t *__restrict__ list,
                                                        int64_t *__restrict__ offsets,
                                                        int32_t *__restrict__ out, int32_t buffer,
                                                        uint32_t *__restrict__ err)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nq)
    return;
  const int64_t q = list ? (int64_t)list[t] : t;
  const RayKeys src{values, ray_pred(rays + 6 * q)};
  int hits = 0;
  for (int j = 0; j < n; ++j)
  {
    const float *v = values + 6 * (int64_t)j;
    hits += ray_distance(src.ray, Box6{{v[0], v[1], v[2]}, {v[3], v[4], v[5]}}) < INFINITY;
  }
  const int m = hits < k ? hits : k;
  int32_t *row;
  if (MODE == MODE_COUNT || (MODE == MODE_BUFFER && m > buffer))
  {
    offsets[q] = m;
    return;
  }
  if (MODE == MODE_BUFFER)
    row = out + q * (int64_t)buffer;
  else
  {
    row = out + offsets[q];
    if (offsets[q + 1] - offsets[q] != m)
    {
      raise_error(err, KERR_LAYOUT);
      return;
    }
  }
  // max-heap of the m smallest keys met so far
  int size = 0;
  uint64_t top = 0;
  for (int j = 0; j < n && m > 0; ++j)
  {
    const uint64_t kj = src.at(j);
    if ((uint32_t)(kj >> 32) >= 0x7f800000u) // +inf: a miss
      continue;
    if (size < m)
    {
      int i = size++;
      while (i > 0)
      {
        const int up = (i - 1) >> 1;
        if (src.at(row[up]) >= kj)
          break;
        row[i] = row[up];
        i = up;
      }
      row[i] = j;
      top = src.at(row[0]);
    }
    else if (kj < top)
    {
      heap_sift_down(src, row, 0, size, j, kj);
      top = src.at(row[0]);
    }
  }
  for (int end = size - 1; end > 0; --end)
  {
    const int32_t v = row[end];
    row[end] = row[0];
    heap_sift_down(src, row, 0, end, v, src.at(v));
  }
  if (MODE == MODE_BUFFER)
    offsets[q] = m;
}

inline unsigned blocks_of(int64_t n) { return (unsigned)div_up(n, 256); }

template <int G, int PK>
hipError_t bf_spatial_gp(const float *values, int64_t n, const float *preds, int64_t nq,
                         float radius, const uint32_t *list, int64_t *offsets, int32_t *out,
                         int32_t buffer, int mode, uint32_t *err, hipStream_t s)
{
  if (mode == MODE_COUNT)
    AXB_LAUNCH((k_bf_spatial<G, PK, MODE_COUNT>), dim3(blocks_of(nq)), dim3(256), 0, s, values,
               n, preds, nq, radius, list, offsets, out, buffer, err);
  else if (mode == MODE_FILL)
    AXB_LAUNCH((k_bf_spatial<G, PK, MODE_FILL>), dim3(blocks_of(nq)), dim3(256), 0, s, values,
               n, preds, nq, radius, list, offsets, out, buffer, err);
  else
    AXB_LAUNCH((k_bf_spatial<G, PK, MODE_BUFFER>), dim3(blocks_of(nq)), dim3(256), 0, s, values,
               n, preds, nq, radius, list, offsets, out, buffer, err);
  return hipGetLastError();
}

template <int G>
hipError_t bf_spatial_g(const float *values, int64_t n, int pred_kind, const float *preds,
                        int64_t nq, float radius, const uint32_t *list, int64_t *offsets,
                        int32_t *out, int32_t buffer, int mode, uint32_t *err, hipStream_t s)
{
  if (pred_kind == PRED_SPHERES)
    return bf_spatial_gp<G, PRED_SPHERES>(values, n, preds, nq, radius, list, offsets, out,
                                          buffer, mode, err, s);
  if (pred_kind == PRED_BOXES)
    return bf_spatial_gp<G, PRED_BOXES>(values, n, preds, nq, radius, list, offsets, out, buffer,
                                        mode, err, s);
  return bf_spatial_gp<G, PRED_RADIUS>(values, n, preds, nq, radius, list, offsets, out, buffer,
                                       mode, err, s);
}

template <int G>
hipError_t bf_nearest_g(const float *values, int n, const float *pts, int64_t nq, int k,
                        int32_t *oi, float *od, hipStream_t s)
{
  const dim3 grid(blocks_of(nq)), block(256);
  if (k <= 1)
    AXB_LAUNCH((k_bf_nearest<G, 1>), grid, block, 0, s, values, n, pts, nq, k, oi, od);
  else if (k <= 2)
    AXB_LAUNCH((k_bf_nearest<G, 2>), grid, block, 0, s, values, n, pts, nq, k, oi, od);
  else if (k <= 4)
    AXB_LAUNCH((k_bf_nearest<G, 4>), grid, block, 0, s, values, n, pts, nq, k, oi, od);
  else if (k <= 8)
    AXB_LAUNCH((k_bf_nearest<G, 8>), grid, block, 0, s, values, n, pts, nq, k, oi, od);
  else if (k <= 16)
    AXB_LAUNCH((k_bf_nearest<G, 16>), grid, block, 0, s, values, n, pts, nq, k, oi, od);
  else if (k <= 32)
    AXB_LAUNCH((k_bf_nearest<G, 32>), grid, block, 0, s, values, n, pts, nq, k, oi, od);
  else if (k <= 64)
    AXB_LAUNCH((k_bf_nearest<G, 64>), grid, block, 0, s, values, n, pts, nq, k, oi, od);
  else
    AXB_LAUNCH((k_bf_heap<G>), grid, block, 0, s, values, n, pts, nq, k, oi, od);
  return hipGetLastError();
}

} // namespace

hipError_t launch_bf_spatial(const float *values, int64_t n, int geometry, int pred_kind,
                             const float *preds, int64_t nq, float radius, const uint32_t *list,
                             int64_t *offsets, int32_t *out, int32_t buffer, int mode,
                             uint32_t *err, hipStream_t s)
{
  if (nq <= 0)
    return hipSuccess;
  if (pred_kind == PRED_RAYS) // boxes only (the caller checks)
    return bf_spatial_gp<GEOM_BOXES, PRED_RAYS>(values, n, preds, nq, radius, list, offsets, out,
                                                buffer, mode, err, s);
  if (geometry == GEOM_POINTS)
    return bf_spatial_g<GEOM_POINTS>(values, n, pred_kind, preds, nq, radius, list, offsets, out,
                                     buffer, mode, err, s);
  return bf_spatial_g<GEOM_BOXES>(values, n, pred_kind, preds, nq, radius, list, offsets, out,
                                  buffer, mode, err, s);
}

hipError_t launch_bf_ray_nearest(const float *values, int64_t n, const float *rays, int64_t nq,
                                 int k, const uint32_t *list, int64_t *offsets, int32_t *out,
                                 int32_t buffer, int mode, uint32_t *err, hipStream_t s)
{
  if (nq <= 0)
    return hipSuccess;
  const dim3 grid(blocks_of(nq)), block(256);
  if (mode == MODE_COUNT)
    AXB_LAUNCH(k_bf_ray_nearest<MODE_COUNT>, grid, block, 0, s, values, (int)n, rays, nq, k,
               list, offsets, out, buffer, err);
  else if (mode == MODE_FILL)
    AXB_LAUNCH(k_bf_ray_nearest<MODE_FILL>, grid, block, 0, s, values, (int)n, rays, nq, k,
               list, offsets, out, buffer, err);
  else
    AXB_LAUNCH(k_bf_ray_nearest<MODE_BUFFER>, grid, block, 0, s, values, (int)n, rays, nq, k,
               list, offsets, out, buffer, err);
  return hipGetLastError();
}

hipError_t launch_bf_nearest(const float *values, int64_t n, int geometry, const float *pts,
                             int64_t nq, int k, int32_t *oi, float *od, hipStream_t s)
{
  if (nq <= 0 || n <= 0 || k <= 0)
    return hipSuccess;
  if (geometry == GEOM_POINTS)
    return bf_nearest_g<GEOM_POINTS>(values, (int)n, pts, nq, k, oi, od, s);
  return bf_nearest_g<GEOM_BOXES>(values, (int)n, pts, nq, k, oi, od, s);
}

} // namespace axb
