// ArborX_Callbacks.hpp -- header-only callback traversals of the MI355X BVH
// (compiled by the caller with hipcc for gfx950; included by ArborX.hpp
// when __HIPCC__ is defined).
//
// Mirrors the generic-callback half of ArborX's query interface:
//   bvh.query(space, predicates, callback[, policy])
//       (src/spatial/ArborX_LinearBVH.hpp:84-88,258-313;
//        detail/ArborX_TreeTraversal.hpp:34-336)
//   bvh.query(space, predicates, callback, values, offsets[, policy])
//       (ArborX_LinearBVH.hpp:90-110; detail/ArborX_CrsGraphWrapperImpl.hpp)
//   bvh.query(PerThread, predicate, callback)  -- device code
//       (ArborX_LinearBVH.hpp:112-122)
// Spatial callbacks run inside the library's stackless rope walk
// (arborx_mi355x_device.hpp) and may return CallbackTreeTraversalControl
// to stop early (detail/ArborX_Callbacks.hpp:96-118).  Nearest callbacks see
// the k nearest values of each predicate in (distance, index) order, taken
// from the library's kNN (TreeTraversal.hpp:314-334 calls them the same way,
// after sorting its heap).
//
// The same two host forms exist for the brute-force index
// (src/spatial/ArborX_BruteForce.hpp:64-91): spatial callbacks see every
// matching value in index order, nearest callbacks as above.
//
// The predicate object handed to a callback carries its geometry and its
// position in the batch: getGeometry(p) and getData(p) (= the position, the
// reference's attach_indices form, detail/ArborX_AttachIndices.hpp:30-35).
// The value is the index's value: the index of the point or box.
#pragma once

#include "ArborX.hpp"
#include "arborx_mi355x_device.hpp"

#include <type_traits>

namespace ArborX
{

// detail/ArborX_Callbacks.hpp:22-26
enum class CallbackTreeTraversalControl
{
  early_exit,
  normal_continuation
};

namespace Experimental
{
// ArborX_LinearBVH.hpp:112-122 tag of the per-thread query.
struct PerThread
{
};
} // namespace Experimental

// Predicate objects seen by callbacks (detail/ArborX_Predicates.hpp:57-146).
template <class Geometry>
struct Intersects
{
  Geometry geometry;
  int index; // position in the batch
};
template <class Geometry>
struct Nearest
{
  Geometry geometry;
  int k;
  int index;
};
// Experimental::OrderedSpatial<G> (detail/ArborX_Predicates.hpp:104-126):
// the callback sees values in increasing (distance(G, value), index) order
// -- for a ray the boxes it hits, for a sphere, box or point every value --
// and may stop the walk with early_exit (first hit).
template <class Geometry>
struct OrderedSpatial
{
  Geometry geometry;
  int index;
};
template <class P>
ARBORX_HD auto const &getGeometry(P const &p)
{
  return p.geometry;
}
template <class P>
ARBORX_HD int getData(P const &p)
{
  return p.index;
}
template <class G>
ARBORX_HD int getK(Nearest<G> const &p)
{
  return p.k;
}

namespace Details
{

// Device view of a tree (the layout of arborx_mi355x_device.hpp); the
// member names are the ones rope_walk reads.
struct DeviceTree
{
  const uint4 *slots = nullptr;
  uint32_t total = 0;
  int32_t n = 0;
  axb::Box6 scene;
  float ss[3];
  uint32_t *err = nullptr;
  int32_t geometry = 0;
  int32_t format = 0;

  // bvh.query(PerThread, predicate, callback) (ArborX_LinearBVH.hpp:112-122)
  template <class Geometry, class Callback>
  __device__ void query(Experimental::PerThread, Intersects<Geometry> const &predicate,
                        Callback const &callback) const;
  template <class Geometry, class Callback>
  __device__ void query(Experimental::PerThread, OrderedSpatial<Geometry> const &predicate,
                        Callback const &callback) const;
};

inline DeviceTree device_tree(axb_bvh const *h)
{
  DeviceTree t;
  if (!h)
    return t;
  axb_tree_view v;
  check(axb_bvh_view(h, &v));
  t.slots = (const uint4 *)v.slots;
  t.total = (uint32_t)v.end_slot;
  t.n = (int32_t)v.n;
  for (int d = 0; d < 3; ++d)
  {
    t.scene.lo[d] = v.scene[d];
    t.scene.hi[d] = v.scene[3 + d];
    t.ss[d] = v.q16_scale[d];
  }
  t.geometry = v.geometry;
  t.format = v.node_format;
  return t;
}

// Calls the callback; false means "stop this predicate's traversal"
// (detail/ArborX_Callbacks.hpp:96-118).
template <class Callback, class Predicate>
__device__ __forceinline__ bool invoke_continue(Callback const &cb, Predicate const &p, int v)
{
  using Result = decltype(cb(p, v)); // evaluated in device context
  if constexpr (std::is_same_v<Result, CallbackTreeTraversalControl>)
    return cb(p, v) != CallbackTreeTraversalControl::early_exit;
  else
  {
    cb(p, v);
    return true;
  }
}

template <class Predicate, class Callback>
struct CallbackSink
{
  Predicate const &p;
  Callback const &cb;
  __device__ __forceinline__ bool hit(int32_t v) { return invoke_continue(cb, p, v); }
};

__device__ __forceinline__ axb::SpatialPred spatial_of(Sphere const &s)
{
  return axb::sphere_pred(s._centroid._coords[0], s._centroid._coords[1],
                          s._centroid._coords[2], s._radius);
}
__device__ __forceinline__ axb::SpatialPred spatial_of(Box const &b)
{
  return axb::box_pred(axb::Box6{
      {b._min_corner._coords[0], b._min_corner._coords[1], b._min_corner._coords[2]},
      {b._max_corner._coords[0], b._max_corner._coords[1], b._max_corner._coords[2]}});
}
__device__ __forceinline__ axb::SpatialPred spatial_of(Experimental::Ray const &r)
{
  return axb::ray_spatial_pred((float const *)&r);
}
template <class Geometry>
constexpr int pred_kind()
{
  if constexpr (std::is_same_v<Geometry, Experimental::Ray>)
    return axb::PRED_RAYS;
  return std::is_same_v<Geometry, Box> ? axb::PRED_BOXES : axb::PRED_SPHERES;
}

// One predicate over the whole tree: the rope walk with the callback as
// its sink.  The branch on the node format and geometry is uniform.
template <class Geometry, class Callback>
__device__ void spatial_walk(DeviceTree const &t, Intersects<Geometry> const &pred,
                             Callback const &cb)
{
  if (t.n <= 0)
    return;
  const axb::SpatialPred sp = spatial_of(pred.geometry);
  CallbackSink<Intersects<Geometry>, Callback> sink{pred, cb};
  uint32_t vi = 0, vl = 0;
  constexpr int PK = pred_kind<Geometry>();
  if constexpr (PK == axb::PRED_RAYS)
  {
    // trees of boxes only (checked on the host)
    if (t.geometry != AXB_GEOMETRY_BOXES)
      return;
    if (t.format == AXB_NODES_Q16)
      axb::rope_walk<axb::NODES_Q16, axb::GEOM_BOXES, PK>(t, sp, 0u, t.total, ~0u, sink, vi,
                                                          vl);
    else
      axb::rope_walk<axb::NODES_FP32, axb::GEOM_BOXES, PK>(t, sp, 0u, t.total, ~0u, sink, vi,
                                                           vl);
    return;
  }
  else if (t.format == AXB_NODES_Q16)
  {
    if (t.geometry == AXB_GEOMETRY_POINTS)
      axb::rope_walk<axb::NODES_Q16, axb::GEOM_POINTS, PK>(t, sp, 0u, t.total, ~0u, sink, vi,
                                                           vl);
    else
      axb::rope_walk<axb::NODES_Q16, axb::GEOM_BOXES, PK>(t, sp, 0u, t.total, ~0u, sink, vi,
                                                          vl);
  }
  else
  {
    if (t.geometry == AXB_GEOMETRY_POINTS)
      axb::rope_walk<axb::NODES_FP32, axb::GEOM_POINTS, PK>(t, sp, 0u, t.total, ~0u, sink, vi,
                                                            vl);
    else
      axb::rope_walk<axb::NODES_FP32, axb::GEOM_BOXES, PK>(t, sp, 0u, t.total, ~0u, sink, vi,
                                                           vl);
  }
}

// The device's form of an ordered predicate's geometry: a ray, or the
// geometry of OrderedSpatial<Sphere | Box | Point>.
__device__ __forceinline__ axb::RayPred ordered_of(Experimental::Ray const &r)
{
  return axb::ray_pred((float const *)&r);
}
__device__ __forceinline__ axb::OrderedGeom ordered_of(Sphere const &s)
{
  const float q[4] = {s._centroid._coords[0], s._centroid._coords[1], s._centroid._coords[2],
                      s._radius};
  return axb::ordered_geom(axb::PRED_SPHERES, q);
}
__device__ __forceinline__ axb::OrderedGeom ordered_of(Box const &b)
{
  const float q[6] = {b._min_corner._coords[0], b._min_corner._coords[1],
                      b._min_corner._coords[2], b._max_corner._coords[0],
                      b._max_corner._coords[1], b._max_corner._coords[2]};
  return axb::ordered_geom(axb::PRED_BOXES, q);
}
__device__ __forceinline__ axb::OrderedGeom ordered_of(Point const &p)
{
  return axb::ordered_geom(axb::PRED_POINTS, p._coords);
}

// One ordered predicate, best-first (TreeTraversal.hpp:338-489): the
// callback gets the values by increasing (distance, index) -- a ray: the
// boxes it hits (trees of boxes only); a sphere, box or point: every value
// -- and may stop the walk.  A walk whose heap of 64 pending entries
// overflows stops and raises KERR_STACK in t.err (when set).
template <class Geometry, class Callback>
__device__ void spatial_walk(DeviceTree const &t, OrderedSpatial<Geometry> const &pred,
                             Callback const &cb)
{
  constexpr bool ray = std::is_same_v<Geometry, Experimental::Ray>;
  if (t.n <= 0 || (ray && t.geometry != AXB_GEOMETRY_BOXES))
    return;
  const auto g = ordered_of(pred.geometry);
  CallbackSink<OrderedSpatial<Geometry>, Callback> sink{pred, cb};
  uint32_t vi = 0, vl = 0;
  if (ray || t.geometry == AXB_GEOMETRY_BOXES)
  {
    if (t.format == AXB_NODES_Q16)
      axb::ordered_walk<axb::NODES_Q16, 64, axb::GEOM_BOXES>(t, g, sink, vi, vl);
    else
      axb::ordered_walk<axb::NODES_FP32, 64, axb::GEOM_BOXES>(t, g, sink, vi, vl);
  }
  else if constexpr (!ray)
  {
    if (t.format == AXB_NODES_Q16)
      axb::ordered_walk<axb::NODES_Q16, 64, axb::GEOM_POINTS>(t, g, sink, vi, vl);
    else
      axb::ordered_walk<axb::NODES_FP32, 64, axb::GEOM_POINTS>(t, g, sink, vi, vl);
  }
}

// Device view of a brute-force index (axb_bruteforce_view).
struct BruteForceView
{
  const float *values = nullptr;
  int64_t n = 0;
  int32_t geometry = 0;
};

inline BruteForceView bruteforce_view(axb_bruteforce const *h)
{
  BruteForceView v;
  if (!h)
    return v;
  axb_values_view c;
  check(axb_bruteforce_view(h, &c));
  v.values = c.values;
  v.n = c.n;
  v.geometry = c.geometry;
  return v;
}

// One predicate against every value of a brute-force index, in index
// order (BruteForceImpl.hpp:53-153), with the same exact leaf test as the
// tree's walk; a callback returning early_exit stops the predicate.
template <class Geometry, class Callback>
__device__ void spatial_walk(BruteForceView const &t, Intersects<Geometry> const &pred,
                             Callback const &cb)
{
  const axb::SpatialPred sp = spatial_of(pred.geometry);
  CallbackSink<Intersects<Geometry>, Callback> sink{pred, cb};
  constexpr int PK = pred_kind<Geometry>();
  if (t.geometry == AXB_GEOMETRY_POINTS)
  {
    // (rays: boxes only, checked on the host)
    if constexpr (PK != axb::PRED_RAYS)
      for (int64_t j = 0; j < t.n; ++j)
      {
        const float *v = t.values + 3 * j;
        const uint4 s0 =
            make_uint4(axb::f_bits(v[0]), axb::f_bits(v[1]), axb::f_bits(v[2]), 0u);
        if (axb::leaf_hit<axb::GEOM_POINTS, PK>(sp, s0, s0) && !sink.hit((int32_t)j))
          return;
      }
  }
  else
  {
    for (int64_t j = 0; j < t.n; ++j)
    {
      const float *v = t.values + 6 * j;
      const uint4 s0 = make_uint4(axb::f_bits(v[0]), axb::f_bits(v[1]), axb::f_bits(v[2]), 0u);
      const uint4 s1 = make_uint4(axb::f_bits(v[3]), axb::f_bits(v[4]), axb::f_bits(v[5]), 0u);
      if (axb::leaf_hit<axb::GEOM_BOXES, PK>(sp, s0, s1) && !sink.hit((int32_t)j))
        return;
    }
  }
}

// Predicate i of a batch as the object a callback sees.
__device__ __forceinline__ Intersects<Sphere> predicate_at(Experimental::PointsWithRadius b,
                                                           int i)
{
  return {Sphere{b.points[i], b.r}, i};
}
__device__ __forceinline__ Intersects<Sphere> predicate_at(Experimental::IntersectsSpheres b,
                                                           int i)
{
  return {b.spheres[i], i};
}
__device__ __forceinline__ Intersects<Box> predicate_at(Experimental::IntersectsBoxes b, int i)
{
  return {b.boxes[i], i};
}
__device__ __forceinline__ Nearest<Point> predicate_at(Experimental::PointsNearestK b, int i)
{
  return {b.points[i], b.k, i};
}
__device__ __forceinline__ Intersects<Experimental::Ray>
predicate_at(Experimental::IntersectsRays b, int i)
{
  return {b.rays[i], i};
}
__device__ __forceinline__ OrderedSpatial<Experimental::Ray>
predicate_at(Experimental::OrderedIntersectsRays b, int i)
{
  return {b.rays[i], i};
}
__device__ __forceinline__ Nearest<Experimental::Ray> predicate_at(Experimental::RaysNearestK b,
                                                                   int i)
{
  return {b.rays[i], b.k, i};
}
__device__ __forceinline__ OrderedSpatial<Sphere>
predicate_at(Experimental::OrderedIntersectsSpheres b, int i)
{
  return {b.spheres[i], i};
}
__device__ __forceinline__ OrderedSpatial<Box> predicate_at(Experimental::OrderedIntersectsBoxes b,
                                                            int i)
{
  return {b.boxes[i], i};
}
__device__ __forceinline__ OrderedSpatial<Point>
predicate_at(Experimental::OrderedIntersectsPoints b, int i)
{
  return {b.points[i], i};
}

inline int32_t c_kind(Experimental::PointsWithRadius const &) { return 2; }
inline int32_t c_kind(Experimental::IntersectsSpheres const &) { return AXB_PREDICATE_SPHERES; }
inline int32_t c_kind(Experimental::IntersectsBoxes const &) { return AXB_PREDICATE_BOXES; }
inline float const *c_data(Experimental::PointsWithRadius const &b)
{
  return (float const *)b.points;
}
inline float const *c_data(Experimental::IntersectsSpheres const &b)
{
  return (float const *)b.spheres;
}
inline float const *c_data(Experimental::IntersectsBoxes const &b)
{
  return (float const *)b.boxes;
}
inline int32_t c_kind(Experimental::IntersectsRays const &) { return AXB_PREDICATE_RAYS; }
inline int32_t c_kind(Experimental::OrderedIntersectsRays const &)
{
  return AXB_PREDICATE_ORDERED_RAYS;
}
inline float const *c_data(Experimental::IntersectsRays const &b)
{
  return (float const *)b.rays;
}
inline float const *c_data(Experimental::OrderedIntersectsRays const &b)
{
  return (float const *)b.rays;
}
inline int32_t c_kind(Experimental::OrderedIntersectsSpheres const &)
{
  return AXB_PREDICATE_ORDERED_SPHERES;
}
inline int32_t c_kind(Experimental::OrderedIntersectsBoxes const &)
{
  return AXB_PREDICATE_ORDERED_BOXES;
}
inline int32_t c_kind(Experimental::OrderedIntersectsPoints const &)
{
  return AXB_PREDICATE_ORDERED_POINTS;
}
inline float const *c_data(Experimental::OrderedIntersectsSpheres const &b)
{
  return (float const *)b.spheres;
}
inline float const *c_data(Experimental::OrderedIntersectsBoxes const &b)
{
  return (float const *)b.boxes;
}
inline float const *c_data(Experimental::OrderedIntersectsPoints const &b)
{
  return (float const *)b.points;
}

template <class Batch>
struct is_ordered_batch
    : std::bool_constant<std::is_same_v<Batch, Experimental::OrderedIntersectsRays> ||
                         std::is_same_v<Batch, Experimental::OrderedIntersectsSpheres> ||
                         std::is_same_v<Batch, Experimental::OrderedIntersectsBoxes> ||
                         std::is_same_v<Batch, Experimental::OrderedIntersectsPoints>>
{
};
template <class Batch>
struct is_spatial_batch
    : std::bool_constant<std::is_same_v<Batch, Experimental::PointsWithRadius> ||
                         std::is_same_v<Batch, Experimental::IntersectsSpheres> ||
                         std::is_same_v<Batch, Experimental::IntersectsBoxes> ||
                         std::is_same_v<Batch, Experimental::IntersectsRays> ||
                         is_ordered_batch<Batch>::value>
{
};
template <class Batch>
struct is_ray_batch
    : std::bool_constant<std::is_same_v<Batch, Experimental::IntersectsRays> ||
                         std::is_same_v<Batch, Experimental::OrderedIntersectsRays> ||
                         std::is_same_v<Batch, Experimental::RaysNearestK>>
{
};
template <class Batch>
struct is_nearest_batch
    : std::bool_constant<std::is_same_v<Batch, Experimental::PointsNearestK> ||
                         std::is_same_v<Batch, Experimental::RaysNearestK>>
{
};

inline unsigned blocks_of(int64_t n) { return (unsigned)((n + 255) / 256); }

// ---- spatial: callback per match
template <class View, class Batch, class Callback>
__global__ __launch_bounds__(256) void k_spatial_callback(View t, Batch batch, int64_t nq,
                                                          const uint32_t *order, Callback cb)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nq)
    return;
  const int i = order ? (int)order[g] : (int)g;
  spatial_walk(t, predicate_at(batch, i), cb);
}

// Counting / writing output functors of the CRS-with-callback form.
struct CountOut
{
  int64_t *c;
  template <class T>
  __device__ void operator()(T const &) const
  {
    ++*c;
  }
};
template <class T>
struct WriteOut
{
  T *out;
  int64_t *c;
  int64_t end; // the row's end: a callback that inserts more than it counted stops there
  __device__ void operator()(T const &v) const
  {
    if (*c < end)
      out[(*c)++] = v;
  }
};

// ---- spatial: CRS with an inserting callback cb(predicate, value, out)
template <class View, class Batch, class Callback, class T, bool FILL>
__global__ __launch_bounds__(256) void k_spatial_insert(View t, Batch batch, int64_t nq,
                                                        const uint32_t *order, Callback cb,
                                                        int64_t *offsets, T *out)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nq)
    return;
  const int i = order ? (int)order[g] : (int)g;
  int64_t c = FILL ? offsets[i] : 0;
  auto pred = predicate_at(batch, i);
  if (FILL)
  {
    WriteOut<T> w{out, &c, offsets[i + 1]};
    spatial_walk(t, pred, [&](auto const &p, int v) { cb(p, v, w); });
  }
  else
  {
    CountOut w{&c};
    spatial_walk(t, pred, [&](auto const &p, int v) { cb(p, v, w); });
    offsets[i] = c;
  }
}

// ---- nearest: the rows of the library's kNN, callback per value in order
template <class Batch, class Callback>
__global__ __launch_bounds__(256) void k_nearest_callback(Batch batch, int64_t nq,
                                                          const int64_t *offsets,
                                                          const int32_t *indices, Callback cb)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq)
    return;
  const auto pred = predicate_at(batch, (int)i);
  for (int64_t j = offsets[i]; j < offsets[i + 1]; ++j)
    cb(pred, indices[j]);
}
template <class Batch, class Callback, class T, bool FILL>
__global__ __launch_bounds__(256) void k_nearest_insert(Batch batch, int64_t nq,
                                                        const int64_t *knn_off,
                                                        const int32_t *knn_ind, Callback cb,
                                                        int64_t *offsets, T *out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq)
    return;
  const auto pred = predicate_at(batch, (int)i);
  int64_t c = FILL ? offsets[i] : 0;
  for (int64_t j = knn_off[i]; j < knn_off[i + 1]; ++j)
  {
    if (FILL)
      cb(pred, knn_ind[j], WriteOut<T>{out, &c, offsets[i + 1]});
    else
      cb(pred, knn_ind[j], CountOut{&c});
  }
  if (!FILL)
    offsets[i] = c;
}

inline void launch_check(const char *what)
{
  hip_check(hipGetLastError(), what);
}

// Processing order of a spatial batch (Morton-sorted unless the policy
// turns sorting off); empty when not sorted.
template <class Batch>
DeviceVector<uint32_t> processing_order(axb_bvh const *h, HipStream const &space,
                                        Batch const &b,
                                        Experimental::TraversalPolicy const &policy)
{
  DeviceVector<uint32_t> order;
  if (policy._sort_predicates && b.n > 1)
  {
    order.resize(b.n);
    check(axb_bvh_predicate_order(h, c_kind(b), c_data(b), (int64_t)b.n, order.data(),
                                  (axb_stream_t)space.stream));
  }
  return order;
}

// The device view and the predicate processing order of an index: the
// tree's walk in Morton order of the predicates (unless the policy turns
// sorting off); the brute-force index in the caller's order.
inline DeviceTree index_view(BoundingVolumeHierarchy const &i)
{
  return device_tree(i.c_handle());
}
inline BruteForceView index_view(BruteForce const &i) { return bruteforce_view(i.c_handle()); }

// Ray predicates need an index of boxes (the reference defines the ray
// tests against boxes only).
template <class Batch, class View>
void check_ray_geometry(View const &t)
{
  if constexpr (is_ray_batch<Batch>::value)
    if (t.n > 0 && t.geometry != AXB_GEOMETRY_BOXES)
      throw std::invalid_argument("query: ray predicates need an index of boxes");
}

// (a template, so that every translation unit may instantiate it)
template <class I>
__global__ void k_zero(I *p, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = 0;
}

// Error word of an ordered walk (heap overflow), read after the launch.
struct WalkErrors
{
  DeviceVector<uint32_t> word;
  void attach(HipStream const &space, DeviceTree &t)
  {
    word.resize(1);
    k_zero<uint32_t><<<1, 64, 0, space.stream>>>(word.data(), 1);
    launch_check("query(callback): error word");
    t.err = word.data();
  }
  void attach(HipStream const &, BruteForceView &) {}
  void check_after(HipStream const &space)
  {
    if (word.size() == 0)
      return;
    uint32_t h = 0;
    hip_check(hipMemcpyAsync(&h, word.data(), sizeof(h), hipMemcpyDeviceToHost, space.stream),
              "hipMemcpyAsync");
    space.fence();
    if (h & axb::KERR_STACK)
      throw SearchException("ordered query: a walk exceeded its heap of 64 entries");
  }
};
template <class Batch>
DeviceVector<uint32_t> index_order(BoundingVolumeHierarchy const &i, HipStream const &space,
                                   Batch const &b, Experimental::TraversalPolicy const &policy)
{
  return processing_order(i.c_handle(), space, b, policy);
}
template <class Batch>
DeviceVector<uint32_t> index_order(BruteForce const &, HipStream const &, Batch const &,
                                   Experimental::TraversalPolicy const &)
{
  return {};
}

template <class Index, class Batch, class Callback>
void callback_query(Index const &index, HipStream const &space, Batch const &b,
                    Callback const &cb, Experimental::TraversalPolicy const &policy)
{
  if (!index.c_handle() || b.n == 0)
    return;
  if constexpr (is_spatial_batch<Batch>::value)
  {
    auto t = index_view(index);
    check_ray_geometry<Batch>(t);
    WalkErrors errs;
    if constexpr (is_ordered_batch<Batch>::value)
      errs.attach(space, t);
    DeviceVector<uint32_t> order = index_order(index, space, b, policy);
    k_spatial_callback<<<blocks_of(b.n), 256, 0, space.stream>>>(t, b, (int64_t)b.n,
                                                                  order.data(), cb);
    launch_check("query(callback)");
    space.fence();
    errs.check_after(space);
  }
  else
  {
    static_assert(is_nearest_batch<Batch>::value, "unsupported predicates");
    DeviceVector<int> ind;
    DeviceVector<int64_t> off;
    index.query(space, b, ind, off, policy);
    k_nearest_callback<Batch, Callback><<<blocks_of(b.n), 256, 0, space.stream>>>(
        b, (int64_t)b.n, off.data(), ind.data(), cb);
    launch_check("query(nearest callback)");
    space.fence();
  }
}

template <class Index, class Batch, class Callback, class T>
void callback_query_crs(Index const &index, HipStream const &space, Batch const &b,
                        Callback const &cb, DeviceVector<T> &values,
                        DeviceVector<int64_t> &offsets,
                        Experimental::TraversalPolicy const &policy)
{
  auto const *h = index.c_handle();
  offsets.resize(b.n + 1);
  k_zero<int64_t><<<blocks_of((int64_t)b.n + 1), 256, 0, space.stream>>>(offsets.data(),
                                                                    (int64_t)b.n + 1);
  launch_check("query(callback, values, offsets)");
  if (!h || b.n == 0)
  {
    space.fence();
    values.resize(0);
    return;
  }
  const int64_t nq = (int64_t)b.n;
  if constexpr (is_spatial_batch<Batch>::value)
  {
    auto t = index_view(index);
    check_ray_geometry<Batch>(t);
    WalkErrors errs;
    if constexpr (is_ordered_batch<Batch>::value)
      errs.attach(space, t);
    DeviceVector<uint32_t> order = index_order(index, space, b, policy);
    k_spatial_insert<decltype(t), Batch, Callback, T, false>
        <<<blocks_of(nq), 256, 0, space.stream>>>(t, b, nq, order.data(), cb, offsets.data(),
                                                   (T *)nullptr);
    launch_check("query(callback, values, offsets): count");
    check(axb_exclusive_scan_i64(offsets.data(), offsets.data(), nq + 1,
                                 (axb_stream_t)space.stream));
    int64_t nnz = 0;
    hip_check(hipMemcpyAsync(&nnz, offsets.data() + nq, sizeof(nnz), hipMemcpyDeviceToHost,
                             space.stream),
              "hipMemcpyAsync");
    space.fence();
    values.resize((size_t)nnz);
    if (nnz > 0)
    {
      k_spatial_insert<decltype(t), Batch, Callback, T, true>
          <<<blocks_of(nq), 256, 0, space.stream>>>(t, b, nq, order.data(), cb, offsets.data(),
                                                     values.data());
      launch_check("query(callback, values, offsets): fill");
    }
    space.fence();
    errs.check_after(space);
  }
  else
  {
    static_assert(is_nearest_batch<Batch>::value, "unsupported predicates");
    DeviceVector<int> ind;
    DeviceVector<int64_t> off;
    index.query(space, b, ind, off, policy);
    k_nearest_insert<Batch, Callback, T, false><<<blocks_of(nq), 256, 0, space.stream>>>(
        b, nq, off.data(), ind.data(), cb, offsets.data(), (T *)nullptr);
    launch_check("query(nearest callback, values, offsets): count");
    check(axb_exclusive_scan_i64(offsets.data(), offsets.data(), nq + 1,
                                 (axb_stream_t)space.stream));
    int64_t nnz = 0;
    hip_check(hipMemcpyAsync(&nnz, offsets.data() + nq, sizeof(nnz), hipMemcpyDeviceToHost,
                             space.stream),
              "hipMemcpyAsync");
    space.fence();
    values.resize((size_t)nnz);
    if (nnz > 0)
    {
      k_nearest_insert<Batch, Callback, T, true><<<blocks_of(nq), 256, 0, space.stream>>>(
          b, nq, off.data(), ind.data(), cb, offsets.data(), values.data());
      launch_check("query(nearest callback, values, offsets): fill");
    }
    space.fence();
  }
}

template <class Geometry, class Callback>
__device__ void DeviceTree::query(Experimental::PerThread, Intersects<Geometry> const &predicate,
                                  Callback const &callback) const
{
  spatial_walk(*this, predicate, callback);
}
template <class Geometry, class Callback>
__device__ void DeviceTree::query(Experimental::PerThread,
                                  OrderedSpatial<Geometry> const &predicate,
                                  Callback const &callback) const
{
  spatial_walk(*this, predicate, callback);
}

} // namespace Details

// query(PerThread, predicate, callback) inside a kernel
// (ArborX_LinearBVH.hpp:112-122): spatial predicates only, as there.
template <class Geometry, class Callback>
__device__ void query(Experimental::PerThread, Details::DeviceTree const &tree,
                      Intersects<Geometry> const &predicate, Callback const &callback)
{
  Details::spatial_walk(tree, predicate, callback);
}
template <class Geometry, class Callback>
__device__ void query(Experimental::PerThread, Details::DeviceTree const &tree,
                      OrderedSpatial<Geometry> const &predicate, Callback const &callback)
{
  Details::spatial_walk(tree, predicate, callback);
}

} // namespace ArborX
