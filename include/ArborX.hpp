// ArborX.hpp -- header-only C++17 facade of the MI355X BVH library.
//
// Mirrors the host interface of ArborX's BVH hot path (src/ArborX.hpp,
// src/spatial/ArborX_LinearBVH.hpp:50-142, ArborX_CrsGraphWrapper.hpp:22-34,
// detail/ArborX_Predicates.hpp, detail/ArborX_PredicateHelpers.hpp,
// detail/ArborX_TraversalPolicy.hpp, misc/ArborX_Exception.hpp) on top of the
// C ABI in arborx_mi355x.h, without Kokkos.  Kokkos Views become
// ArborX::DeviceVector (device memory owned by the facade); the execution
// space becomes ArborX::HipStream.  Values are points or boxes in device
// memory; a leaf's value is its index (the reference's benchmark form
// BVH<MemorySpace, int> over Iota, benchmarks/bvh_driver/
// benchmark_registration.hpp:151-161).
//
//   ArborX::HipStream space;
//   ArborX::DeviceVector<ArborX::Point> points = ...;
//   ArborX::BoundingVolumeHierarchy bvh(space, points);
//   ArborX::DeviceVector<int> indices;  ArborX::DeviceVector<int64_t> offsets;
//   bvh.query(space, ArborX::Experimental::make_intersects(points, r), indices, offsets);
//   ArborX::query(bvh, space, ArborX::Experimental::make_nearest(points, k), indices, offsets);
#pragma once

#include "arborx_mi355x.h"

#include <hip/hip_runtime_api.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

// Geometry helpers callable from device code when compiled by hipcc (the
// callback traversals of ArborX_Callbacks.hpp); plain host functions under
// a host compiler.
#if defined(__HIPCC__)
#define ARBORX_HD __host__ __device__ inline
#else
#define ARBORX_HD inline
#endif

namespace ArborX
{

// misc/ArborX_Exception.hpp:19-37
class SearchException : public std::logic_error
{
public:
  using std::logic_error::logic_error;
};

namespace Details
{
inline void check(axb_status s)
{
  if (s == AXB_OK)
    return;
  const std::string msg = axb_last_error();
  if (s == AXB_ERR_BUFFER_OVERFLOW)
    throw SearchException(msg);
  throw std::runtime_error("arborx_mi355x: " + msg);
}
inline void hip_check(hipError_t e, const char *what)
{
  if (e != hipSuccess)
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
} // namespace Details

// Execution space: one HIP stream (the default stream unless given one).
struct HipStream
{
  hipStream_t stream = nullptr;
  void fence() const { Details::hip_check(hipStreamSynchronize(stream), "fence"); }
};

// geometry/ArborX_Point.hpp, ArborX_Box.hpp (Box.hpp:34-44 empty box),
// ArborX_Sphere.hpp -- 3-D float, same bytes as the C ABI arrays.
struct Point
{
  float _coords[3];
  ARBORX_HD float &operator[](int d) { return _coords[d]; }
  ARBORX_HD float const &operator[](int d) const { return _coords[d]; }
};
struct Box
{
  Point _min_corner{{FLT_MAX, FLT_MAX, FLT_MAX}};
  Point _max_corner{{-FLT_MAX, -FLT_MAX, -FLT_MAX}};
  Box() = default;
  ARBORX_HD Box(Point const &lo, Point const &hi) : _min_corner(lo), _max_corner(hi) {}
  ARBORX_HD Point &minCorner() { return _min_corner; }
  ARBORX_HD Point const &minCorner() const { return _min_corner; }
  ARBORX_HD Point &maxCorner() { return _max_corner; }
  ARBORX_HD Point const &maxCorner() const { return _max_corner; }
};
struct Sphere
{
  Point _centroid;
  float _radius;
  ARBORX_HD Point const &centroid() const { return _centroid; }
  ARBORX_HD float radius() const { return _radius; }
};

// distance(Point, Point) (geometry/algorithms/ArborX_Distance.hpp:52-70):
// squared differences summed in axis order, then sqrt.
ARBORX_HD float distance(Point const &a, Point const &b)
{
  float s = 0.f;
  for (int d = 0; d < 3; ++d)
  {
    float t = a._coords[d] - b._coords[d];
    s += t * t;
  }
  return std::sqrt(s);
}
static_assert(sizeof(Point) == 12 && sizeof(Box) == 24 && sizeof(Sphere) == 16,
              "facade geometry must match the C ABI layout");

namespace Experimental
{
// geometry/ArborX_Ray.hpp:28-66: origin and unit direction.  The
// constructor normalises the direction with the norm and the division in
// double (misc/ArborX_Vector.hpp:41-55,77-87); a zero component is stored
// as +0 (the library's slab test gives -0 no meaning of its own).  Same
// bytes as a row of the C ABI's ray arrays.
struct Ray
{
  Point _origin = {};
  Point _direction = {};
  Ray() = default;
  ARBORX_HD Ray(Point const &origin, Point const &direction) : _origin(origin)
  {
    double r = 0;
    for (int d = 0; d < 3; ++d)
      r += static_cast<double>(direction[d]) * static_cast<double>(direction[d]);
    const double mag = std::sqrt(r);
    for (int d = 0; d < 3; ++d)
    {
      const float w = static_cast<float>(direction[d] / mag);
      _direction[d] = w == 0.f ? 0.f : w;
    }
  }
  ARBORX_HD Point const &origin() const { return _origin; }
  ARBORX_HD Point const &direction() const { return _direction; }
};
static_assert(sizeof(Ray) == 24, "facade rays must match the C ABI layout");
} // namespace Experimental

// Owning device array (stands in for Kokkos::View<T*, HIP memory space>).
template <class T>
class DeviceVector
{
public:
  DeviceVector() = default;
  explicit DeviceVector(size_t n) { resize(n); }
  explicit DeviceVector(std::vector<T> const &host) { assign(host); }
  DeviceVector(DeviceVector const &) = delete;
  DeviceVector &operator=(DeviceVector const &) = delete;
  DeviceVector(DeviceVector &&o) noexcept { swap(o); }
  DeviceVector &operator=(DeviceVector &&o) noexcept
  {
    swap(o);
    return *this;
  }
  ~DeviceVector() { release(); }

  void resize(size_t n)
  {
    release();
    if (n)
      Details::hip_check(hipMalloc((void **)&_p, n * sizeof(T)), "hipMalloc");
    _n = n;
  }
  void assign(std::vector<T> const &host)
  {
    resize(host.size());
    if (_n)
      Details::hip_check(hipMemcpy(_p, host.data(), _n * sizeof(T), hipMemcpyHostToDevice),
                         "hipMemcpy");
  }
  std::vector<T> to_host() const
  {
    std::vector<T> h(_n);
    if (_n)
      Details::hip_check(hipMemcpy(h.data(), _p, _n * sizeof(T), hipMemcpyDeviceToHost),
                         "hipMemcpy");
    return h;
  }
  // Takes ownership of memory the C ABI allocated (hipMallocAsync).
  void adopt(T *p, size_t n, hipStream_t s)
  {
    release();
    _p = p;
    _n = n;
    _async = true;
    _stream = s;
  }
  T *data() const { return _p; }
  size_t size() const { return _n; }
  size_t extent(int) const { return _n; }

private:
  void release()
  {
    if (_p)
    {
      if (_async)
        (void)axb_free(_p, (axb_stream_t)_stream);
      else
        (void)hipFree(_p);
    }
    _p = nullptr;
    _n = 0;
    _async = false;
  }
  void swap(DeviceVector &o) noexcept
  {
    std::swap(_p, o._p);
    std::swap(_n, o._n);
    std::swap(_async, o._async);
    std::swap(_stream, o._stream);
  }
  T *_p = nullptr;
  size_t _n = 0;
  bool _async = false;
  hipStream_t _stream = nullptr;
};

// ArborX_PairValueIndex.hpp:22-46: a value with its position in the values
// the index was built from (the values of a tree built over
// attach_indices(values)).  32-bit indices.
template <class Value, class Index = unsigned>
struct PairValueIndex
{
  static_assert(std::is_integral_v<Index> && sizeof(Index) == 4, "32-bit indices");
  using value_type = Value;
  using index_type = Index;

  Value value;
  Index index;
};

namespace Experimental
{
// ArborX_AttachIndices.hpp:23-65: values whose positions travel with them;
// a tree built over them returns PairValueIndex<Value, Index> values.
template <class Values, class Index>
struct AttachIndices
{
  Values const &_values;
};
template <class Index = unsigned, class Values>
AttachIndices<Values, Index> attach_indices(Values const &values)
{
  return AttachIndices<Values, Index>{values};
}
} // namespace Experimental

namespace Experimental
{
// detail/ArborX_TraversalPolicy.hpp:19-47
struct TraversalPolicy
{
  int _buffer_size = 0;
  bool _sort_predicates = true;
  TraversalPolicy &setBufferSize(int buffer_size)
  {
    _buffer_size = buffer_size;
    return *this;
  }
  TraversalPolicy &setPredicateSorting(bool sort_predicates)
  {
    _sort_predicates = sort_predicates;
    return *this;
  }
};

// Predicate batches (the results of make_intersects / make_nearest,
// detail/ArborX_PredicateHelpers.hpp:55-206).
struct PointsWithRadius
{
  Point const *points;
  size_t n;
  float r;
};
struct PointsNearestK
{
  Point const *points;
  size_t n;
  int k;
};
struct IntersectsSpheres
{
  Sphere const *spheres;
  size_t n;
};
struct IntersectsBoxes
{
  Box const *boxes;
  size_t n;
};
// Intersects<Ray>, OrderedSpatial<Ray> (ordered_intersects) and
// Nearest<Ray> batches over box trees (geometry/ArborX_Ray.hpp,
// detail/ArborX_Predicates.hpp:104-155).
struct IntersectsRays
{
  Ray const *rays;
  size_t n;
};
struct OrderedIntersectsRays
{
  Ray const *rays;
  size_t n;
};
struct RaysNearestK
{
  Ray const *rays;
  size_t n;
  int k;
};
// OrderedSpatial<Sphere | Box | Point> batches (ordered_intersects of a
// sphere, a box or a point, any tree): values by increasing
// (distance(geometry, value), index) -- every value, since all distances
// are finite.
struct OrderedIntersectsSpheres
{
  Sphere const *spheres;
  size_t n;
};
struct OrderedIntersectsBoxes
{
  Box const *boxes;
  size_t n;
};
struct OrderedIntersectsPoints
{
  Point const *points;
  size_t n;
};

// make_intersects(points, r): spheres of radius r (stored as float),
// PredicateHelpers.hpp:99-105 (r > 0 asserted at :102).
inline PointsWithRadius make_intersects(DeviceVector<Point> const &points, float r)
{
  if (!(r > 0))
    throw std::invalid_argument("make_intersects: r must be > 0");
  return {points.data(), points.size(), r};
}
inline PointsNearestK make_nearest(DeviceVector<Point> const &points, int k)
{
  return {points.data(), points.size(), k};
}
inline RaysNearestK make_nearest(DeviceVector<Ray> const &rays, int k)
{
  return {rays.data(), rays.size(), k};
}
// ordered_intersects / make_ordered_intersects over a batch of rays
// (detail/ArborX_Predicates.hpp:148-155).
inline OrderedIntersectsRays ordered_intersects(DeviceVector<Ray> const &rays)
{
  return {rays.data(), rays.size()};
}
inline OrderedIntersectsRays make_ordered_intersects(DeviceVector<Ray> const &rays)
{
  return {rays.data(), rays.size()};
}
inline OrderedIntersectsSpheres ordered_intersects(DeviceVector<Sphere> const &s)
{
  return {s.data(), s.size()};
}
inline OrderedIntersectsBoxes ordered_intersects(DeviceVector<Box> const &b)
{
  return {b.data(), b.size()};
}
inline OrderedIntersectsPoints ordered_intersects(DeviceVector<Point> const &p)
{
  return {p.data(), p.size()};
}
template <class T>
inline auto make_ordered_intersects(DeviceVector<T> const &g)
    -> decltype(ordered_intersects(g))
{
  return ordered_intersects(g);
}
} // namespace Experimental

// intersects(geometry) / nearest(geometry, k) over a batch
// (detail/ArborX_Predicates.hpp:129-140).
inline Experimental::IntersectsSpheres intersects(DeviceVector<Sphere> const &s)
{
  return {s.data(), s.size()};
}
inline Experimental::IntersectsBoxes intersects(DeviceVector<Box> const &b)
{
  return {b.data(), b.size()};
}
inline Experimental::PointsNearestK nearest(DeviceVector<Point> const &p, int k = 1)
{
  return {p.data(), p.size(), k};
}
inline Experimental::IntersectsRays intersects(DeviceVector<Experimental::Ray> const &r)
{
  return {r.data(), r.size()};
}
inline Experimental::RaysNearestK nearest(DeviceVector<Experimental::Ray> const &r, int k = 1)
{
  return {r.data(), r.size(), k};
}

template <class T>
class DeviceVector;
template <class T>
struct is_device_vector : std::false_type
{
};
template <class T>
struct is_device_vector<DeviceVector<T>> : std::true_type
{
};

#if defined(__HIPCC__)
namespace Details
{
struct DeviceTree;
DeviceTree device_tree(axb_bvh const *h);
// Index: BoundingVolumeHierarchy or BruteForce
template <class Index, class Batch, class Callback>
void callback_query(Index const &index, HipStream const &space, Batch const &b,
                    Callback const &cb, Experimental::TraversalPolicy const &policy);
template <class Index, class Batch, class Callback, class T>
void callback_query_crs(Index const &index, HipStream const &space, Batch const &b,
                        Callback const &cb, DeviceVector<T> &values,
                        DeviceVector<int64_t> &offsets,
                        Experimental::TraversalPolicy const &policy);
} // namespace Details
#endif

enum class NodeFormat
{
  FP32 = AXB_NODES_FP32,
  Q16 = AXB_NODES_Q16
};

// spatial/ArborX_LinearBVH.hpp:50-142.
class BoundingVolumeHierarchy
{
public:
  using size_type = int64_t;
  using bounding_volume_type = Box;

  BoundingVolumeHierarchy() = default; // empty tree

  BoundingVolumeHierarchy(HipStream const &space, DeviceVector<Point> const &values,
                          NodeFormat fmt = NodeFormat::FP32)
  {
    build(space, (float const *)values.data(), (int64_t)values.size(), AXB_GEOMETRY_POINTS,
          fmt);
  }
  BoundingVolumeHierarchy(HipStream const &space, DeviceVector<Box> const &values,
                          NodeFormat fmt = NodeFormat::FP32)
  {
    build(space, (float const *)values.data(), (int64_t)values.size(), AXB_GEOMETRY_BOXES,
          fmt);
  }
  // Built over attach_indices(values) (ArborX_LinearBVH.hpp:50-57 with
  // PairValueIndex values): the tree keeps a copy of the values, and the
  // query overload taking DeviceVector<PairValueIndex<...>> returns them
  // with their indices.
  template <class G, class Index>
  BoundingVolumeHierarchy(HipStream const &space,
                          Experimental::AttachIndices<DeviceVector<G>, Index> const &values,
                          NodeFormat fmt = NodeFormat::FP32)
      : BoundingVolumeHierarchy(space, values._values, fmt)
  {
    static_assert(std::is_same_v<G, Point> || std::is_same_v<G, Box>, "points or boxes");
    DeviceVector<G> const &v = values._values;
    _values.resize(v.size() * (sizeof(G) / sizeof(float)));
    if (v.size())
      Details::hip_check(hipMemcpyAsync(_values.data(), v.data(), v.size() * sizeof(G),
                                        hipMemcpyDeviceToDevice, space.stream),
                         "hipMemcpyAsync");
    _value_floats = (int)(sizeof(G) / sizeof(float));
  }
  BoundingVolumeHierarchy(BoundingVolumeHierarchy const &) = delete;
  BoundingVolumeHierarchy &operator=(BoundingVolumeHierarchy const &) = delete;
  BoundingVolumeHierarchy(BoundingVolumeHierarchy &&o) noexcept { swap(o); }
  BoundingVolumeHierarchy &operator=(BoundingVolumeHierarchy &&o) noexcept
  {
    swap(o);
    return *this;
  }
  ~BoundingVolumeHierarchy()
  {
    if (_h)
      axb_bvh_destroy(_h);
  }

  size_type size() const noexcept { return _h ? axb_bvh_size(_h) : 0; }
  bool empty() const noexcept { return size() == 0; }
  Box bounds() const
  {
    Box b;
    if (_h)
    {
      float f[6];
      Details::check(axb_bvh_bounds(_h, f));
      b = Box({{f[0], f[1], f[2]}}, {{f[3], f[4], f[5]}});
    }
    return b;
  }

  // query(space, predicates, indices, offsets[, policy])
  // (ArborX_LinearBVH.hpp:90-110): indices/offsets are (re)allocated.
  void query(HipStream const &space, Experimental::PointsWithRadius const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    prepare(offsets, preds.n);
    int32_t *ind = nullptr;
    int64_t nnz = 0;
    if (_h)
      Details::check(axb_bvh_query_intersects_radius(
          _h, (float const *)preds.points, (int64_t)preds.n, preds.r, c_policy(policy),
          offsets.data(), &ind, &nnz, (axb_stream_t)space.stream));
    finish(space, indices, ind, nnz, offsets, preds.n);
  }
  void query(HipStream const &space, Experimental::IntersectsSpheres const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    prepare(offsets, preds.n);
    int32_t *ind = nullptr;
    int64_t nnz = 0;
    if (_h)
      Details::check(axb_bvh_query_intersects(
          _h, AXB_PREDICATE_SPHERES, (float const *)preds.spheres, (int64_t)preds.n,
          c_policy(policy), offsets.data(), &ind, &nnz, (axb_stream_t)space.stream));
    finish(space, indices, ind, nnz, offsets, preds.n);
  }
  void query(HipStream const &space, Experimental::IntersectsBoxes const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    prepare(offsets, preds.n);
    int32_t *ind = nullptr;
    int64_t nnz = 0;
    if (_h)
      Details::check(axb_bvh_query_intersects(
          _h, AXB_PREDICATE_BOXES, (float const *)preds.boxes, (int64_t)preds.n,
          c_policy(policy), offsets.data(), &ind, &nnz, (axb_stream_t)space.stream));
    finish(space, indices, ind, nnz, offsets, preds.n);
  }
  // Rays over a tree of boxes: every box the ray meets (rows in traversal
  // order), the same boxes by increasing (distance, index), or the first k
  // of those (rows may be shorter than k).
  void query(HipStream const &space, Experimental::IntersectsRays const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    rays_impl(space, AXB_PREDICATE_RAYS, preds.rays, preds.n, -1, indices, offsets, policy);
  }
  void query(HipStream const &space, Experimental::OrderedIntersectsRays const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    rays_impl(space, AXB_PREDICATE_ORDERED_RAYS, preds.rays, preds.n, -1, indices, offsets,
              policy);
  }
  void query(HipStream const &space, Experimental::RaysNearestK const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    rays_impl(space, AXB_PREDICATE_ORDERED_RAYS, preds.rays, preds.n, preds.k, indices, offsets,
              policy);
  }
  // ordered_intersects of spheres, boxes or points: every value of the
  // tree per row, by (distance, index).
  void query(HipStream const &space, Experimental::OrderedIntersectsSpheres const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    rays_impl(space, AXB_PREDICATE_ORDERED_SPHERES, preds.spheres, preds.n, -1, indices,
              offsets, policy);
  }
  void query(HipStream const &space, Experimental::OrderedIntersectsBoxes const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    rays_impl(space, AXB_PREDICATE_ORDERED_BOXES, preds.boxes, preds.n, -1, indices, offsets,
              policy);
  }
  void query(HipStream const &space, Experimental::OrderedIntersectsPoints const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    rays_impl(space, AXB_PREDICATE_ORDERED_POINTS, preds.points, preds.n, -1, indices, offsets,
              policy);
  }
  void query(HipStream const &space, Experimental::PointsNearestK const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    DeviceVector<float> *no_dist = nullptr;
    nearest_impl(space, preds, indices, offsets, no_dist, policy);
  }
  // Nearest with the float distances of every result (same row layout).
  void query(HipStream const &space, Experimental::PointsNearestK const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             DeviceVector<float> &distances,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    nearest_impl(space, preds, indices, offsets, &distances, policy);
  }

  // query(space, predicates, values, offsets[, policy]) for a tree built
  // over attach_indices: each result as PairValueIndex{value, index}, rows
  // as the index form's.
  template <class Predicates, class V, class I>
  void query(HipStream const &space, Predicates const &preds,
             DeviceVector<PairValueIndex<V, I>> &values, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    static_assert(sizeof(PairValueIndex<V, I>) == sizeof(V) + 4 && sizeof(V) % 4 == 0,
                  "packed (value, index) records");
    if (!empty() && _value_floats != (int)(sizeof(V) / sizeof(float)))
      throw std::invalid_argument(
          "BVH::query: the tree was not built over attach_indices of values of this type");
    DeviceVector<int> idx;
    query(space, preds, idx, offsets, policy);
    values.resize(idx.size());
    if (idx.size())
      Details::check(axb_gather_pairs(_values.data(), _value_floats, size(), idx.data(),
                                      (int64_t)idx.size(), values.data(),
                                      (axb_stream_t)space.stream));
  }

  axb_bvh const *c_handle() const { return _h; }

#if defined(__HIPCC__)
  // query(space, predicates, callback[, policy]) (ArborX_LinearBVH.hpp:84-88):
  // callback(predicate, value) per match; spatial callbacks may return
  // CallbackTreeTraversalControl::early_exit.  Synchronises the stream.
  template <class Predicates, class Callback,
            std::enable_if_t<!is_device_vector<Callback>::value, int> = 0>
  void query(HipStream const &space, Predicates const &predicates, Callback const &callback,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    Details::callback_query(*this, space, predicates, callback, policy);
  }
  // query(space, predicates, callback, values, offsets[, policy])
  // (ArborX_LinearBVH.hpp:90-110): callback(predicate, value, out) calls
  // out(x) for each element x of the predicate's row.
  template <class Predicates, class Callback, class T>
  void query(HipStream const &space, Predicates const &predicates, Callback const &callback,
             DeviceVector<T> &values, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    Details::callback_query_crs(*this, space, predicates, callback, values, offsets, policy);
  }
  // Device view for query(PerThread, ...) inside the caller's kernels;
  // valid while the tree lives.
  Details::DeviceTree device_view() const;
#endif

private:
  static axb_traversal_policy c_policy(Experimental::TraversalPolicy const &p)
  {
    return axb_traversal_policy{p._buffer_size, p._sort_predicates ? 1 : 0};
  }
  void build(HipStream const &space, float const *v, int64_t n, int32_t geometry,
             NodeFormat fmt)
  {
    Details::check(axb_bvh_create(v, n, geometry, (int32_t)fmt, 0,
                                  (axb_stream_t)space.stream, &_h));
  }
  static void prepare(DeviceVector<int64_t> &offsets, size_t nq)
  {
    if (offsets.size() != nq + 1)
      offsets.resize(nq + 1);
  }
  void finish(HipStream const &space, DeviceVector<int> &indices, int32_t *ind, int64_t nnz,
              DeviceVector<int64_t> &offsets, size_t nq) const
  {
    if (!_h)
    {
      // empty tree: every row is empty
      std::vector<int64_t> z(nq + 1, 0);
      offsets.assign(z);
      indices.resize(0);
      return;
    }
    indices.adopt(ind, (size_t)nnz, space.stream);
  }
  // kind: AXB_PREDICATE_RAYS or one of the ordered kinds; k >= 0:
  // nearest(ray, k)
  template <class G>
  void rays_impl(HipStream const &space, int32_t kind, G const *rays, size_t n, int k,
                 DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
                 Experimental::TraversalPolicy const &policy) const
  {
    prepare(offsets, n);
    int32_t *ind = nullptr;
    int64_t nnz = 0;
    if (_h && k >= 0)
      Details::check(axb_bvh_query_nearest_rays(_h, (float const *)rays, (int64_t)n, k,
                                                c_policy(policy), offsets.data(), &ind, &nnz,
                                                (axb_stream_t)space.stream));
    else if (_h)
      Details::check(axb_bvh_query_intersects(_h, kind, (float const *)rays, (int64_t)n,
                                              c_policy(policy), offsets.data(), &ind, &nnz,
                                              (axb_stream_t)space.stream));
    finish(space, indices, ind, nnz, offsets, n);
  }
  void nearest_impl(HipStream const &space, Experimental::PointsNearestK const &preds,
                    DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
                    DeviceVector<float> *distances,
                    Experimental::TraversalPolicy const &policy) const
  {
    prepare(offsets, preds.n);
    if (!_h)
    {
      finish(space, indices, nullptr, 0, offsets, preds.n);
      if (distances)
        distances->resize(0);
      return;
    }
    int32_t *ind = nullptr;
    float *dist = nullptr;
    int64_t nnz = 0;
    Details::check(axb_bvh_query_nearest(_h, (float const *)preds.points, (int64_t)preds.n,
                                         preds.k, c_policy(policy), offsets.data(), &ind,
                                         distances ? &dist : nullptr, &nnz,
                                         (axb_stream_t)space.stream));
    indices.adopt(ind, (size_t)nnz, space.stream);
    if (distances)
      distances->adopt(dist, (size_t)nnz, space.stream);
  }

  void swap(BoundingVolumeHierarchy &o) noexcept
  {
    std::swap(_h, o._h);
    std::swap(_values, o._values);
    std::swap(_value_floats, o._value_floats);
  }

  axb_bvh *_h = nullptr;
  DeviceVector<float> _values; // attach_indices: a copy of the values
  int _value_floats = 0;       // 3 (points) or 6 (boxes); 0 without them
};

using BVH = BoundingVolumeHierarchy;

// spatial/ArborX_BruteForce.hpp:40-178: the BVH's construction and CRS query
// interface over an index that tests every predicate against every value.
// Spatial rows list values in increasing index order; nearest rows are
// ordered by (distance, index).
class BruteForce
{
public:
  using size_type = int64_t;
  using bounding_volume_type = Box;

  BruteForce() = default; // empty index

  BruteForce(HipStream const &space, DeviceVector<Point> const &values)
  {
    build(space, (float const *)values.data(), (int64_t)values.size(), AXB_GEOMETRY_POINTS);
  }
  BruteForce(HipStream const &space, DeviceVector<Box> const &values)
  {
    build(space, (float const *)values.data(), (int64_t)values.size(), AXB_GEOMETRY_BOXES);
  }
  BruteForce(BruteForce const &) = delete;
  BruteForce &operator=(BruteForce const &) = delete;
  BruteForce(BruteForce &&o) noexcept { std::swap(_h, o._h); }
  BruteForce &operator=(BruteForce &&o) noexcept
  {
    std::swap(_h, o._h);
    return *this;
  }
  ~BruteForce()
  {
    if (_h)
      axb_bruteforce_destroy(_h);
  }

  size_type size() const noexcept { return _h ? axb_bruteforce_size(_h) : 0; }
  bool empty() const noexcept { return size() == 0; }
  Box bounds() const
  {
    Box b;
    if (_h)
    {
      float f[6];
      Details::check(axb_bruteforce_bounds(_h, f));
      b = Box({{f[0], f[1], f[2]}}, {{f[3], f[4], f[5]}});
    }
    return b;
  }

  // query(space, predicates, indices, offsets[, policy])
  // (ArborX_BruteForce.hpp:71-91): indices/offsets are (re)allocated.
  void query(HipStream const &space, Experimental::PointsWithRadius const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    prepare(offsets, preds.n);
    int32_t *ind = nullptr;
    int64_t nnz = 0;
    if (_h)
      Details::check(axb_bruteforce_query_intersects_radius(
          _h, (float const *)preds.points, (int64_t)preds.n, preds.r, c_policy(policy),
          offsets.data(), &ind, &nnz, (axb_stream_t)space.stream));
    finish(space, indices, ind, nnz, offsets, preds.n);
  }
  void query(HipStream const &space, Experimental::IntersectsSpheres const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    prepare(offsets, preds.n);
    int32_t *ind = nullptr;
    int64_t nnz = 0;
    if (_h)
      Details::check(axb_bruteforce_query_intersects(
          _h, AXB_PREDICATE_SPHERES, (float const *)preds.spheres, (int64_t)preds.n,
          c_policy(policy), offsets.data(), &ind, &nnz, (axb_stream_t)space.stream));
    finish(space, indices, ind, nnz, offsets, preds.n);
  }
  void query(HipStream const &space, Experimental::IntersectsBoxes const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    prepare(offsets, preds.n);
    int32_t *ind = nullptr;
    int64_t nnz = 0;
    if (_h)
      Details::check(axb_bruteforce_query_intersects(
          _h, AXB_PREDICATE_BOXES, (float const *)preds.boxes, (int64_t)preds.n,
          c_policy(policy), offsets.data(), &ind, &nnz, (axb_stream_t)space.stream));
    finish(space, indices, ind, nnz, offsets, preds.n);
  }
  // nearest(ray, k) over boxes: rows of at most k values by (distance, index).
  void query(HipStream const &space, Experimental::RaysNearestK const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    prepare(offsets, preds.n);
    int32_t *ind = nullptr;
    int64_t nnz = 0;
    if (_h)
      Details::check(axb_bruteforce_query_nearest_rays(
          _h, (float const *)preds.rays, (int64_t)preds.n, preds.k, c_policy(policy),
          offsets.data(), &ind, &nnz, (axb_stream_t)space.stream));
    finish(space, indices, ind, nnz, offsets, preds.n);
  }
  // Intersects<Ray> over boxes (the reference's brute force has no ordered
  // traversal).
  void query(HipStream const &space, Experimental::IntersectsRays const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    prepare(offsets, preds.n);
    int32_t *ind = nullptr;
    int64_t nnz = 0;
    if (_h)
      Details::check(axb_bruteforce_query_intersects(
          _h, AXB_PREDICATE_RAYS, (float const *)preds.rays, (int64_t)preds.n,
          c_policy(policy), offsets.data(), &ind, &nnz, (axb_stream_t)space.stream));
    finish(space, indices, ind, nnz, offsets, preds.n);
  }
  void query(HipStream const &space, Experimental::PointsNearestK const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    nearest_impl(space, preds, indices, offsets, nullptr, policy);
  }
  void query(HipStream const &space, Experimental::PointsNearestK const &preds,
             DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
             DeviceVector<float> &distances,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    nearest_impl(space, preds, indices, offsets, &distances, policy);
  }

  axb_bruteforce const *c_handle() const { return _h; }

#if defined(__HIPCC__)
  // query(space, predicates, callback[, policy]) (ArborX_BruteForce.hpp:
  // 64-67,151-178): callback(predicate, value) per match, values in index
  // order; nearest callbacks in (distance, index) order.  Synchronises.
  template <class Predicates, class Callback,
            std::enable_if_t<!is_device_vector<Callback>::value, int> = 0>
  void query(HipStream const &space, Predicates const &predicates, Callback const &callback,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    Details::callback_query(*this, space, predicates, callback, policy);
  }
  // query(space, predicates, callback, values, offsets[, policy])
  // (ArborX_BruteForce.hpp:69-91): callback(predicate, value, out).
  template <class Predicates, class Callback, class T>
  void query(HipStream const &space, Predicates const &predicates, Callback const &callback,
             DeviceVector<T> &values, DeviceVector<int64_t> &offsets,
             Experimental::TraversalPolicy const &policy = {}) const
  {
    Details::callback_query_crs(*this, space, predicates, callback, values, offsets, policy);
  }
#endif

private:
  static axb_traversal_policy c_policy(Experimental::TraversalPolicy const &p)
  {
    return axb_traversal_policy{p._buffer_size, p._sort_predicates ? 1 : 0};
  }
  void build(HipStream const &space, float const *v, int64_t n, int32_t geometry)
  {
    Details::check(axb_bruteforce_create(v, n, geometry, (axb_stream_t)space.stream, &_h));
  }
  static void prepare(DeviceVector<int64_t> &offsets, size_t nq)
  {
    if (offsets.size() != nq + 1)
      offsets.resize(nq + 1);
  }
  void finish(HipStream const &space, DeviceVector<int> &indices, int32_t *ind, int64_t nnz,
              DeviceVector<int64_t> &offsets, size_t nq) const
  {
    if (!_h)
    {
      std::vector<int64_t> z(nq + 1, 0);
      offsets.assign(z);
      indices.resize(0);
      return;
    }
    indices.adopt(ind, (size_t)nnz, space.stream);
  }
  void nearest_impl(HipStream const &space, Experimental::PointsNearestK const &preds,
                    DeviceVector<int> &indices, DeviceVector<int64_t> &offsets,
                    DeviceVector<float> *distances,
                    Experimental::TraversalPolicy const &policy) const
  {
    prepare(offsets, preds.n);
    if (!_h)
    {
      finish(space, indices, nullptr, 0, offsets, preds.n);
      if (distances)
        distances->resize(0);
      return;
    }
    int32_t *ind = nullptr;
    float *dist = nullptr;
    int64_t nnz = 0;
    Details::check(axb_bruteforce_query_nearest(
        _h, (float const *)preds.points, (int64_t)preds.n, preds.k, c_policy(policy),
        offsets.data(), &ind, distances ? &dist : nullptr, &nnz, (axb_stream_t)space.stream));
    indices.adopt(ind, (size_t)nnz, space.stream);
    if (distances)
      distances->adopt(dist, (size_t)nnz, space.stream);
  }

  axb_bruteforce *_h = nullptr;
};

namespace Details
{
inline void neighbor_list(HipStream const &space, DeviceVector<Point> const &points,
                          float radius, int32_t full, DeviceVector<int64_t> &offsets,
                          DeviceVector<int> &indices)
{
  if (offsets.size() != points.size() + 1)
    offsets.resize(points.size() + 1);
  int32_t *ind = nullptr;
  int64_t nnz = 0;
  check(axb_find_neighbor_list((float const *)points.data(), (int64_t)points.size(), radius,
                               full, offsets.data(), &ind, &nnz, (axb_stream_t)space.stream));
  indices.adopt(ind, (size_t)nnz, space.stream);
}

// detail/ArborX_ExpandHalfToFull.hpp:23-68: replaces a half list by the
// full one (offsets and indices are reallocated, as there).
inline void expandHalfToFull(HipStream const &space, DeviceVector<int64_t> &offsets,
                             DeviceVector<int> &indices)
{
  if (offsets.size() == 0)
    return;
  DeviceVector<int64_t> out(offsets.size());
  int32_t *ind = nullptr;
  int64_t nnz = 0;
  check(axb_expand_half_to_full(offsets.data(), indices.data(), (int64_t)offsets.size() - 1,
                                out.data(), &ind, &nnz, (axb_stream_t)space.stream));
  offsets = std::move(out);
  indices.adopt(ind, (size_t)nnz, space.stream);
}
} // namespace Details

namespace Experimental
{
// detail/ArborX_NeighborList.hpp:105-188: row i lists every j != i within
// `radius` of point i.
inline void findFullNeighborList(HipStream const &space, DeviceVector<Point> const &points,
                                 float radius, DeviceVector<int64_t> &offsets,
                                 DeviceVector<int> &indices)
{
  Details::neighbor_list(space, points, radius, 1, offsets, indices);
}
// detail/ArborX_NeighborList.hpp:49-103: each pair once, in the row of the
// point whose leaf comes later in the tree.
inline void findHalfNeighborList(HipStream const &space, DeviceVector<Point> const &points,
                                 float radius, DeviceVector<int64_t> &offsets,
                                 DeviceVector<int> &indices)
{
  Details::neighbor_list(space, points, radius, 0, offsets, indices);
}
} // namespace Experimental

// Free function ArborX::query(tree, space, predicates, indices, offsets
// [, policy]) (src/spatial/ArborX_CrsGraphWrapper.hpp:22-34).
template <class Predicates, class... Args>
void query(BoundingVolumeHierarchy const &tree, HipStream const &space,
           Predicates const &predicates, Args &&...args)
{
  tree.query(space, predicates, std::forward<Args>(args)...);
}
template <class Predicates, class... Args>
void query(BruteForce const &index, HipStream const &space, Predicates const &predicates,
           Args &&...args)
{
  index.query(space, predicates, std::forward<Args>(args)...);
}

} // namespace ArborX

#if defined(__HIPCC__)
#include "ArborX_Callbacks.hpp"
inline ArborX::Details::DeviceTree ArborX::BoundingVolumeHierarchy::device_view() const
{
  return Details::device_tree(_h);
}
#endif

