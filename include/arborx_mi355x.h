/* arborx_mi355x.h -- C-ABI of the MI355X (gfx950) BVH library.
 *
 * Plain C: pointers, sizes and status codes only.  Every `values`,
 * `predicates`, `points`, `offsets`, `indices` and `distances` pointer is a
 * DEVICE pointer; arrays of 3-D points are N x 3 float (AoS, x y z), arrays of
 * boxes are N x 6 float (min x y z, max x y z), spheres are N x 4 float
 * (centre x y z, radius).  All work is enqueued on `stream` (a hipStream_t;
 * NULL = the null stream).  Functions return AXB_OK (0) or a negative
 * axb_status; axb_last_error() gives the message of the calling thread's last
 * failure.
 *
 * Each entry point names the ArborX interface it replaces (paths relative to
 * the ArborX source tree).  The reference's Kokkos Views are replaced by
 * device pointers; views the reference reallocates itself
 * (CrsGraphWrapperImpl.hpp:376-377) are allocated here through the output
 * allocator (axb_set_output_allocator, default hipMallocAsync on `stream`)
 * and released with axb_free().
 */
#ifndef ARBORX_MI355X_H
#define ARBORX_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same type as hipStream_t, declared here so that callers need no HIP
 * headers. */
typedef struct ihipStream_t *axb_stream_t;

/* Opaque tree handle; replaces ArborX::BoundingVolumeHierarchy
 * (src/spatial/ArborX_LinearBVH.hpp:50-142). */
typedef struct axb_bvh axb_bvh;

typedef enum axb_status
{
  AXB_OK = 0,
  AXB_ERR_INVALID_ARGUMENT = -1,
  AXB_ERR_HIP = -2,
  AXB_ERR_OUT_OF_MEMORY = -3,
  /* A query produced more results than a hard (negative) buffer_size allows.
   * Replaces ArborX::SearchException thrown by ARBORX_ASSERT at
   * src/spatial/detail/ArborX_CrsGraphWrapperImpl.hpp:268. */
  AXB_ERR_BUFFER_OVERFLOW = -4,
  /* A kernel reported an internal limit (bounded wait exceeded, kNN stack
   * overflow). */
  AXB_ERR_INTERNAL = -5
} axb_status;

enum
{
  AXB_GEOMETRY_POINTS = 0, /* values are N x 3 points */
  AXB_GEOMETRY_BOXES = 1   /* values are N x 6 boxes */
};

enum
{
  AXB_NODES_FP32 = 0, /* internal node boxes as 6 x float (32 B / node) */
  AXB_NODES_Q16 = 1   /* internal node boxes as 6 x uint16, outward-rounded
                         relative to a frame box (16 B / node) */
};

enum
{
  AXB_BUILD_KEEP_TOPOLOGY = 1 /* keep the reference-form arrays (sorted Morton
                                 codes, permutation, left_child, rope, boxes)
                                 for axb_bvh_export_topology */
};

enum
{
  AXB_PREDICATE_SPHERES = 0, /* Intersects<Sphere>: nq x 4 {x,y,z,r} */
  AXB_PREDICATE_BOXES = 1,   /* Intersects<Box>:    nq x 6 {lo, hi} */
  /* Rays, nq x 6 {origin xyz, unit direction xyz} (Experimental::Ray,
   * src/geometry/ArborX_Ray.hpp:28-66; axb_make_rays builds them), box
   * trees only.  Intersects<Ray>: the boxes the ray's half-line meets
   * (ArborX_Ray.hpp:156-167). */
  AXB_PREDICATE_RAYS = 3,
  /* OrderedSpatial<Ray> (Experimental::ordered_intersects,
   * detail/ArborX_Predicates.hpp:104-155, TreeTraversal.hpp:338-489): the
   * same boxes, each row in increasing (distance, index) order, distance =
   * the ray parameter where it enters the box (ArborX_Ray.hpp:429-444).  The
   * reference leaves equal distances in the order of its traversal heap. */
  AXB_PREDICATE_ORDERED_RAYS = 4,
  /* OrderedSpatial<Sphere | Box | Point> (ordered_intersects of a sphere
   * {x,y,z,r}, nq x 4; a box {lo, hi}, nq x 6; a point, nq x 3), any tree:
   * the ordered traversal reports every value at a finite distance
   * (TreeTraversal.hpp:338-489 pushes every child below +inf), so each row
   * lists all n values in increasing (distance, index) order, distance =
   * distance(geometry, value) (geometry/algorithms/ArborX_Distance.hpp:
   * 52-208).  Rows are n long: meant for small trees (the callback form with
   * early exit is the common use). */
  AXB_PREDICATE_ORDERED_SPHERES = 5,
  AXB_PREDICATE_ORDERED_BOXES = 6,
  AXB_PREDICATE_ORDERED_POINTS = 7
};

/* ArborX::Experimental::TraversalPolicy
 * (src/spatial/detail/ArborX_TraversalPolicy.hpp:19-47).
 * buffer_size 0: two passes (count, fill); > 0: one pass into a buffer of
 * buffer_size results per query, second pass only if some query overflows;
 * < 0: like > 0 but overflow is an error (AXB_ERR_BUFFER_OVERFLOW).
 * sort_predicates != 0: process queries in Morton-32 order of their centroid
 * (results are still returned in the caller's query order). */
typedef struct axb_traversal_policy
{
  int32_t buffer_size;
  int32_t sort_predicates;
} axb_traversal_policy;

typedef void *(*axb_alloc_fn)(size_t bytes, void *ctx);
typedef void (*axb_free_fn)(void *ptr, void *ctx);

const char *axb_last_error(void);
const char *axb_version(void);

/* Allocator for the result arrays the library sizes itself (CRS indices,
 * distances).  alloc/free NULL restores the default (hipMallocAsync /
 * hipFreeAsync on the query stream). */
axb_status axb_set_output_allocator(axb_alloc_fn alloc, axb_free_fn free_fn,
                                    void *ctx);
/* Releases an array returned by a query when the default allocator is in use. */
axb_status axb_free(void *device_ptr, axb_stream_t stream);
/* The library's temporaries come from the current device's default
 * stream-ordered memory pool, which the library sets to keep freed memory
 * cached (returning it to the driver at every synchronisation costs
 * milliseconds per call).  This returns the cached memory of the current
 * device's default pool to the driver. */
axb_status axb_trim_memory(void);

/* Node array layout: an internal node whose subtree holds at most this many
 * leaves is not stored; its leaves follow each other in the pre-order array
 * (the walk tests each of them, so results are unchanged). */
#define AXB_FLAT_SUBTREE_MAX 4

/* Build: BoundingVolumeHierarchy(space, values[, indexable_getter, Morton64])
 * (src/spatial/ArborX_LinearBVH.hpp:68-73,167-256).  Values are points or
 * boxes; leaf i's value is its index in `values` (the reference's benchmark
 * form BVH<MemorySpace, int, ...> over Iota / attach_indices,
 * benchmarks/bvh_driver/benchmark_registration.hpp:151-161).
 * flags: AXB_BUILD_KEEP_TOPOLOGY or 0.  n must be < 2^30.  The caller keeps
 * ownership of `values`; the tree copies what it needs. */
axb_status axb_bvh_create(const float *values, int64_t n, int32_t geometry,
                          int32_t node_format, int32_t flags,
                          axb_stream_t stream, axb_bvh **out);
void axb_bvh_destroy(axb_bvh *bvh);
/* size() / empty() (ArborX_LinearBVH.hpp:75-79). */
int64_t axb_bvh_size(const axb_bvh *bvh);
/* bounds() (ArborX_LinearBVH.hpp:81-82): host copy {lo xyz, hi xyz}; an
 * empty tree returns the empty box {+FLT_MAX, -FLT_MAX}. */
axb_status axb_bvh_bounds(const axb_bvh *bvh, float out_host[6]);

/* Spatial query to CRS: bvh.query(space, predicates, indices, offsets,
 * policy) with Intersects<Sphere> or Intersects<Box> predicates
 * (ArborX_LinearBVH.hpp:90-110, ArborX_CrsGraphWrapper.hpp:22-34,
 * detail/ArborX_CrsGraphWrapperImpl.hpp:376-446).
 * offsets: caller-allocated, nq + 1 entries (int64; the reference uses the
 * offset View's type).  *indices: allocated by the library, *nnz_host
 * entries, each the index of a value whose geometry satisfies the
 * predicate; row i lists them in traversal order. */
axb_status axb_bvh_query_intersects(const axb_bvh *bvh, int32_t predicate_kind,
                                    const float *predicates, int64_t nq,
                                    axb_traversal_policy policy,
                                    int64_t *offsets, int32_t **indices,
                                    int64_t *nnz_host, axb_stream_t stream);

/* Ray construction, Experimental::Ray(origin, direction)
 * (src/geometry/ArborX_Ray.hpp:42-51): rays (device, n x 6) receive
 * {origins[i], directions[i] / |directions[i]|}, the norm and the division
 * in double (misc/ArborX_Vector.hpp:41-55,77-87).  A zero component is
 * stored as +0.  A zero direction is an error (the reference asserts
 * norm > 0).  origins, directions: device n x 3. */
axb_status axb_make_rays(const float *origins, const float *directions, int64_t n,
                         float *rays, axb_stream_t stream);

/* Nearest<Ray> to CRS: nearest(ray, k) over a box tree
 * (detail/ArborX_TreeTraversal.hpp:122-336 with distance(Ray, Box),
 * ArborX_Ray.hpp:429-444): row i holds the first min(k, hits) boxes of ray
 * i in (distance, index) order -- boxes the ray misses are at +inf and never
 * qualify, so rows may be shorter than k.  Arguments as for
 * axb_bvh_query_intersects. */
axb_status axb_bvh_query_nearest_rays(const axb_bvh *bvh, const float *rays, int64_t nq,
                                      int32_t k, axb_traversal_policy policy,
                                      int64_t *offsets, int32_t **indices,
                                      int64_t *nnz_host, axb_stream_t stream);

/* make_intersects(points, r) (detail/ArborX_PredicateHelpers.hpp:99-105,
 * 163-186): spheres of one radius r (stored as float) around nq points. */
axb_status axb_bvh_query_intersects_radius(const axb_bvh *bvh,
                                           const float *points, int64_t nq,
                                           float radius,
                                           axb_traversal_policy policy,
                                           int64_t *offsets, int32_t **indices,
                                           int64_t *nnz_host,
                                           axb_stream_t stream);

/* Nearest query to CRS: bvh.query(space, make_nearest(points, k), indices,
 * offsets) (detail/ArborX_PredicateHelpers.hpp:114-119,188-206,
 * CrsGraphWrapperImpl.hpp:353-374, TreeTraversal.hpp:122-336).  Row i holds
 * min(k, size()) indices ordered by (distance, index); *distances (optional,
 * pass NULL to skip) receives the matching float distances.  offsets:
 * caller-allocated nq + 1 entries. */
axb_status axb_bvh_query_nearest(const axb_bvh *bvh, const float *points,
                                 int64_t nq, int32_t k,
                                 axb_traversal_policy policy, int64_t *offsets,
                                 int32_t **indices, float **distances,
                                 int64_t *nnz_host, axb_stream_t stream);

/* Neighbour lists of a point set with itself,
 * Experimental::findFullNeighborList / findHalfNeighborList
 * (src/spatial/detail/ArborX_NeighborList.hpp:49-188): row i lists every
 * j != i with distance(x_i, x_j) <= radius (full), or only the j whose leaf
 * precedes i's leaf in the tree (half; each pair appears once, as in
 * detail/ArborX_HalfTraversal.hpp:45-72).  axb_bvh_neighbor_list runs on a
 * points tree built by axb_bvh_create (rows indexed by value); the find_
 * form builds a Q16 tree over `points` first, as the reference builds its
 * own BVH.  offsets: caller-allocated n + 1 entries; *indices allocated by
 * the library (*nnz_host entries), each row in traversal order (the
 * reference's order within a row is unspecified).  policy as for queries;
 * radius >= 0. */
axb_status axb_bvh_neighbor_list(const axb_bvh *bvh, float radius, int32_t full,
                                 axb_traversal_policy policy, int64_t *offsets,
                                 int32_t **indices, int64_t *nnz_host,
                                 axb_stream_t stream);
axb_status axb_find_neighbor_list(const float *points, int64_t n, float radius,
                                  int32_t full, int64_t *offsets,
                                  int32_t **indices, int64_t *nnz_host,
                                  axb_stream_t stream);

/* Details::expandHalfToFull (detail/ArborX_ExpandHalfToFull.hpp:23-68):
 * from a half list (n rows) to the full list, row i = its own entries
 * entries plus the k of every row k that lists i (order within a row
 * unspecified, as in the reference).  out_offsets: caller-allocated n + 1;
 * *out_indices allocated by the library, 2 * nnz(half) entries. */
axb_status axb_expand_half_to_full(const int64_t *offsets,
                                   const int32_t *indices, int64_t n,
                                   int64_t *out_offsets, int32_t **out_indices,
                                   int64_t *nnz_host, axb_stream_t stream);

/* ---- brute-force index ----
 * ArborX::BruteForce<MemorySpace, Value> (src/spatial/ArborX_BruteForce.hpp:
 * 40-178, detail/ArborX_BruteForceImpl.hpp:34-229): the same construction
 * and query interface as the BVH, answered by testing every predicate
 * against every value -- an alternative index for small sets and an
 * independent check of the tree.  The index keeps a device copy of the
 * values.  Spatial rows list the values in increasing index order (the
 * reference's order inside a row is unspecified); nearest rows are ordered
 * by (distance, index) as for the BVH.  policy.buffer_size as for the BVH
 * queries; sort_predicates has no effect. */
typedef struct axb_bruteforce axb_bruteforce;
/* BruteForce(space, values) (ArborX_BruteForce.hpp:52-53,117-150); n < 2^31. */
axb_status axb_bruteforce_create(const float *values, int64_t n,
                                 int32_t geometry, axb_stream_t stream,
                                 axb_bruteforce **out);
void axb_bruteforce_destroy(axb_bruteforce *bf);
/* size() / bounds() (ArborX_BruteForce.hpp:56-62); bounds of an empty index
 * are {+FLT_MAX, -FLT_MAX}. */
int64_t axb_bruteforce_size(const axb_bruteforce *bf);
axb_status axb_bruteforce_bounds(const axb_bruteforce *bf, float out_host[6]);
/* query(space, predicates, indices, offsets[, policy])
 * (ArborX_BruteForce.hpp:66-91,151-178, BruteForceImpl.hpp:53-153): as for
 * axb_bvh_query_intersects / _radius; rows in increasing value index.
 * AXB_PREDICATE_RAYS over boxes is accepted; AXB_PREDICATE_ORDERED_RAYS is
 * not (the reference's brute force has no ordered traversal; nearest rays:
 * axb_bruteforce_query_nearest_rays). */
axb_status axb_bruteforce_query_intersects(const axb_bruteforce *bf,
                                           int32_t predicate_kind,
                                           const float *predicates, int64_t nq,
                                           axb_traversal_policy policy,
                                           int64_t *offsets, int32_t **indices,
                                           int64_t *nnz_host,
                                           axb_stream_t stream);
axb_status axb_bruteforce_query_intersects_radius(
    const axb_bruteforce *bf, const float *points, int64_t nq, float radius,
    axb_traversal_policy policy, int64_t *offsets, int32_t **indices,
    int64_t *nnz_host, axb_stream_t stream);
/* Nearest (BruteForceImpl.hpp:155-229): arguments as for
 * axb_bvh_query_nearest. */
axb_status axb_bruteforce_query_nearest(const axb_bruteforce *bf,
                                        const float *points, int64_t nq,
                                        int32_t k, axb_traversal_policy policy,
                                        int64_t *offsets, int32_t **indices,
                                        float **distances, int64_t *nnz_host,
                                        axb_stream_t stream);

/* Nearest<Ray> over a brute-force index of boxes (BruteForceImpl.hpp:155-229
 * with distance(Ray, Box)): as axb_bvh_query_nearest_rays. */
axb_status axb_bruteforce_query_nearest_rays(const axb_bruteforce *bf, const float *rays,
                                             int64_t nq, int32_t k, axb_traversal_policy policy,
                                             int64_t *offsets, int32_t **indices,
                                             int64_t *nnz_host, axb_stream_t stream);

/* Read-only view of a brute-force index for device code (the callback
 * forms of query over it, include/ArborX_Callbacks.hpp): the index's device
 * copy of its values (n x 3 or n x 6 floats) and their geometry.  Valid
 * while the index lives. */
typedef struct axb_values_view
{
  const float *values;
  int64_t n;
  int32_t geometry;
} axb_values_view;
axb_status axb_bruteforce_view(const axb_bruteforce *bf, axb_values_view *out_host);

/* Pair energy over a fixed neighbour CRS (no reference counterpart; pattern of
 * examples/molecular_dynamics/example_molecular_dynamics.cpp:128-194).
 * Shifted Lennard-Jones phi(r) = 4(r^-12 - r^-6) - phi(cutoff);
 * E = 1/2 sum_i sum_{j in row i, j != i} phi(|x_i - x_j|), pair terms in
 * float, sums in double.  energy (device double, may be NULL) receives E;
 * grad (device N x 3 float, may be NULL) receives dE/dx. */
axb_status axb_pair_energy(const float *xyz, int64_t n, const int64_t *offsets,
                           const int32_t *indices, float cutoff,
                           double *energy, float *grad, axb_stream_t stream);

/* The same energy and gradient without any allocation (no reference
 * counterpart).  For callers that cannot allocate device memory: code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc.  Its
 * scratch (the partial sums of the energy kernel's blocks, at most 2^19) is
 * a static device array of 2^19 doubles (4 MiB) that every GPU the library
 * runs on keeps until the process ends; n above
 * AXB_PAIR_ENERGY_NOALLOC_MAX_N is rejected with AXB_ERR_INVALID_ARGUMENT.
 * Rows are taken in the caller's order (axb_pair_energy may reorder large
 * inputs, which changes E only by the summation order of its partials; the
 * gradient is bit-identical).  Calls that overlap on one GPU share its
 * scratch and race: serialise them (one stream, or events). */
#define AXB_PAIR_ENERGY_NOALLOC_MAX_N (INT64_C(1) << 27)
axb_status axb_pair_energy_noalloc(const float *xyz, int64_t n, const int64_t *offsets,
                                   const int32_t *indices, float cutoff,
                                   double *energy, float *grad, axb_stream_t stream);

/* Central finite differences of the same energy in double precision for the
 * atoms listed in `atoms` (device int32, m entries): fd[3*a + d] =
 * (E(x + h e_ad) - E(x - h e_ad)) / 2h with the neighbour set held fixed
 * (device double, m x 3). */
axb_status axb_pair_energy_fd(const float *xyz, int64_t n,
                              const int64_t *offsets, const int32_t *indices,
                              float cutoff, const int32_t *atoms, int64_t m,
                              double h, double *fd, axb_stream_t stream);

/* ---- building blocks of the header-only callback traversals ----
 * (query(space, predicates, callback) and query(PerThread, ...) in
 * include/ArborX.hpp, compiled by the caller with hipcc; they walk the
 * tree's node array with the device code of include/arborx_mi355x_device.hpp).
 *
 * Read-only view of a tree for device code.  slots: pre-order node array
 * (uint32 x 4 per slot, walked from slot 0 to end_slot); first_leaf_prefix:
 * n entries locating each leaf's record; scene: {lo xyz, hi xyz};
 * q16_scale: per-axis step of the scene frame (Q16 trees).  Valid while the
 * tree lives. */
typedef struct axb_tree_view
{
  const void *slots;
  int64_t end_slot;
  int64_t n;
  const int32_t *first_leaf_prefix;
  float scene[6];
  float q16_scale[3];
  int32_t geometry;
  int32_t node_format;
} axb_tree_view;
axb_status axb_bvh_view(const axb_bvh *bvh, axb_tree_view *out_host);

/* Processing order of a predicate batch: the permutation that sorts the
 * predicates' centroids by 32-bit Morton code over the tree's scene
 * (CrsGraphWrapperImpl.hpp:407-419, the reference's sortPredicates).
 * predicate_kind: AXB_PREDICATE_SPHERES / _BOXES, 2 for points (3
 * floats each), or AXB_PREDICATE_RAYS / _ORDERED_RAYS (a ray's centroid is
 * its origin, ArborX_Ray.hpp:76-80).  order: device uint32, nq entries
 * (caller-allocated), nq < 2^30 (the sort's counts are 30-bit; queries
 * with sort_predicates set take the same limit). */
axb_status axb_bvh_predicate_order(const axb_bvh *bvh, int32_t predicate_kind,
                                   const float *predicates, int64_t nq,
                                   uint32_t *order, axb_stream_t stream);

/* Results as (value, index) records: the PairValueIndex values of a tree
 * built over attach_indices(values) (ArborX_PairValueIndex.hpp:22-46,
 * ArborX_AttachIndices.hpp:23-65), for the C++ facade's query overload
 * that returns them.  Record j of out (device, nnz records of
 * 4 * (value_floats + 1) bytes) = {values[indices[j]] (value_floats floats:
 * 3 for points, 6 for boxes), indices[j] (32-bit)}.  values: device,
 * n_values x value_floats floats; an index outside [0, n_values) fails with
 * AXB_ERR_INVALID_ARGUMENT.  Synchronises the stream. */
axb_status axb_gather_pairs(const float *values, int32_t value_floats, int64_t n_values,
                            const int32_t *indices, int64_t nnz, void *out,
                            axb_stream_t stream);

/* Exclusive prefix sum of n int64 values (device; in place allowed),
 * the scan of the two-pass CRS (KokkosExt::exclusive_scan). */
axb_status axb_exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n,
                                  axb_stream_t stream);

/* ---- introspection (parity tests, benchmarks) ---- */

/* Number of 16-byte slots of the pre-order node array and a device pointer
 * to it (uint32 x 4 per slot; read-only, owned by the tree). */
int64_t axb_bvh_num_slots(const axb_bvh *bvh);
const void *axb_bvh_slots(const axb_bvh *bvh);

/* Copies the reference-form topology (tree built with
 * AXB_BUILD_KEEP_TOPOLOGY): sorted Morton codes (n), permutation (n), leaf
 * ropes (n), left_child (n-1), internal ropes (n-1), internal boxes
 * ((n-1) x 6).  Node indices follow detail/ArborX_HappyTreeFriends.hpp:23-86
 * (leaves 0..n-1, internal node k is n+k, rope -1 = sentinel).  Device
 * destination pointers; any may be NULL. */
axb_status axb_bvh_export_topology(const axb_bvh *bvh, uint64_t *keys,
                                   uint32_t *perm, int32_t *leaf_rope,
                                   int32_t *left_child, int32_t *rope,
                                   float *boxes, axb_stream_t stream);

/* Q16 encoding of one box against a frame (host; the function the build's
 * encoder runs): q = {qlo xyz, qhi xyz}, the tightest outward-rounded steps
 * (see the node layout).  For tests. */
void axb_q16_encode_host(const float frame[6], const float box[6], uint16_t q[6]);

/* ---- profiling (replaces the Kokkos Tools labels the reference puts on
 * every region and kernel, spatial/ArborX_LinearBVH.hpp:201,217-255) ----
 * flags bit 0: bracket every kernel launch with HIP events on its stream;
 * bit 1: count the node records each traversal reads.  0 disables.
 * Enabling clears previous records. */
axb_status axb_profile_enable(int32_t flags);
axb_status axb_profile_clear(void);
int32_t axb_profile_num_records(void);
/* Kernel name and duration (ms) of record i (waits for its events). */
axb_status axb_profile_record(int32_t i, char *name, int32_t name_len,
                              double *ms);
/* {queries, internal-node records read, leaf records read, results} summed
 * over the traversals run since axb_profile_enable(2 | ...), then the same
 * four counters for the nearest queries that took the exact fallback walk,
 * then {collect rounds run, queries needing more than one round, 0, 0}. */
axb_status axb_profile_visits(int64_t out_host[12]);

/* Test hook (no reference counterpart): lowers limits the library normally
 * derives from the device, so that small inputs run the paths large ones
 * take -- kNN queries per collect launch (several launches and one
 * fallback launch after them), bytes of the scratch rows of a buffered
 * CRS query (below the rows' need: plain count + fill), and collect rounds
 * before a kNN query goes to the exact fallback walk (-1: none, every query
 * takes the fallback).  0 restores a default.  Process-wide; set it between
 * calls. */
axb_status axb_debug_set_limits(int64_t knn_batch_max, int64_t scratch_rows_bytes,
                                int32_t knn_rounds);

/* Host helper used by the query kernels, exported for unit tests: the largest
 * float s with RN(sqrt(s)) <= r, so that `d2 <= t` decides exactly the
 * reference's `sqrt(d2) <= r` (detail/../geometry/algorithms/
 * ArborX_Intersects.hpp:105-114 with ArborX_Distance.hpp:52-70). */
float axb_sqrt_threshold(float r);

/* Morton-64 code of one point in a scene box (host, for unit tests;
 * detail/ArborX_MortonCode.hpp:318-347 with TranslateAndScale.hpp:23-37). */
uint64_t axb_morton64_host(const float p[3], const float scene[6]);

#ifdef __cplusplus
}
#endif

#endif /* ARBORX_MI355X_H */
