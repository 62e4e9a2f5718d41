"""Argument checks of the C ABI, on the CPU: every call below is rejected
before it touches the GPU, with AXB_ERR_INVALID_ARGUMENT and a message in
axb_last_error() (the reference's ARBORX_ASSERT / KOKKOS_ASSERT
preconditions: PredicateHelpers.hpp:102 r > 0, LinearBVH.hpp size limits,
a null tree or output)."""
import ctypes

import arborx_amd as ax

P = ctypes.c_void_p
I64 = ctypes.c_int64
INVALID = ax.AXB_ERR_INVALID_ARGUMENT


def _last_error():
    return ax.lib().axb_last_error().decode()


def _policy():
    return ax.TraversalPolicy()._c()


def test_create_rejects_bad_arguments():
    L = ax.lib()
    h = P()
    dummy = P(16)  # never dereferenced: the checks come first
    assert L.axb_bvh_create(dummy, -1, ax.GEOMETRY_POINTS, ax.NODES_Q16, 0, None,
                            ctypes.byref(h)) == INVALID
    assert "n must be" in _last_error()
    assert L.axb_bvh_create(dummy, 1 << 30, ax.GEOMETRY_POINTS, ax.NODES_Q16, 0, None,
                            ctypes.byref(h)) == INVALID
    assert L.axb_bvh_create(dummy, 10, 7, ax.NODES_Q16, 0, None, ctypes.byref(h)) == INVALID
    assert "geometry" in _last_error()
    assert L.axb_bvh_create(dummy, 10, ax.GEOMETRY_POINTS, 9, 0, None,
                            ctypes.byref(h)) == INVALID
    assert "node format" in _last_error()
    assert L.axb_bvh_create(None, 10, ax.GEOMETRY_POINTS, ax.NODES_Q16, 0, None,
                            ctypes.byref(h)) == INVALID
    assert L.axb_bvh_create(dummy, 10, ax.GEOMETRY_POINTS, ax.NODES_Q16, 0, None,
                            None) == INVALID
    assert L.axb_bruteforce_create(dummy, -5, ax.GEOMETRY_BOXES, None,
                                   ctypes.byref(h)) == INVALID
    assert L.axb_bruteforce_create(dummy, 5, 3, None, ctypes.byref(h)) == INVALID
    assert L.axb_bruteforce_create(None, 5, ax.GEOMETRY_BOXES, None,
                                   ctypes.byref(h)) == INVALID


def test_queries_reject_bad_arguments():
    L = ax.lib()
    ind = P()
    dist = P()
    nnz = I64()
    dummy = P(16)
    # make_intersects(points, r) requires r > 0 (PredicateHelpers.hpp:102)
    for r in (0.0, -1.0, float("nan")):
        assert L.axb_bvh_query_intersects_radius(dummy, dummy, 4, r, _policy(), dummy,
                                                 ctypes.byref(ind), ctypes.byref(nnz),
                                                 None) == INVALID
        assert L.axb_bruteforce_query_intersects_radius(dummy, dummy, 4, r, _policy(), dummy,
                                                        ctypes.byref(ind), ctypes.byref(nnz),
                                                        None) == INVALID
    assert "radius" in _last_error()
    # unknown predicate kinds; the brute-force index has no ordered traversal
    for kind in (-1, 2, 8, 42):
        assert L.axb_bvh_query_intersects(dummy, kind, dummy, 4, _policy(), dummy,
                                          ctypes.byref(ind), ctypes.byref(nnz), None) == INVALID
    for kind in (-1, 2, 4, 5, 6, 7, 8):
        assert L.axb_bruteforce_query_intersects(dummy, kind, dummy, 4, _policy(), dummy,
                                                 ctypes.byref(ind), ctypes.byref(nnz),
                                                 None) == INVALID
    # null tree / outputs, query counts out of range
    assert L.axb_bvh_query_intersects(None, ax.PREDICATE_SPHERES, dummy, 4, _policy(), dummy,
                                      ctypes.byref(ind), ctypes.byref(nnz), None) == INVALID
    assert L.axb_bvh_query_nearest(None, dummy, 4, 3, _policy(), dummy, ctypes.byref(ind),
                                   ctypes.byref(dist), ctypes.byref(nnz), None) == INVALID
    assert L.axb_bvh_query_nearest(dummy, dummy, -1, 3, _policy(), dummy, ctypes.byref(ind),
                                   ctypes.byref(dist), ctypes.byref(nnz), None) == INVALID
    assert L.axb_bruteforce_query_nearest(None, dummy, 4, 3, _policy(), dummy,
                                          ctypes.byref(ind), ctypes.byref(dist),
                                          ctypes.byref(nnz), None) == INVALID
    assert L.axb_bruteforce_query_nearest(dummy, dummy, 1 << 31, 3, _policy(), dummy,
                                          ctypes.byref(ind), ctypes.byref(dist),
                                          ctypes.byref(nnz), None) == INVALID
    assert "nq must be" in _last_error()


def test_ray_arguments():
    L = ax.lib()
    ind = P()
    nnz = I64()
    dummy = P(16)
    assert L.axb_make_rays(dummy, dummy, -1, dummy, None) == INVALID
    assert L.axb_make_rays(None, dummy, 3, dummy, None) == INVALID
    assert L.axb_make_rays(dummy, dummy, 0, dummy, None) == ax.AXB_OK
    # kinds past the ordered kinds, and the ordered kind on the brute force
    assert L.axb_bvh_query_intersects(dummy, 8, dummy, 4, _policy(), dummy, ctypes.byref(ind),
                                      ctypes.byref(nnz), None) == INVALID
    assert L.axb_bruteforce_query_intersects(dummy, ax.PREDICATE_ORDERED_RAYS, dummy, 4,
                                             _policy(), dummy, ctypes.byref(ind),
                                             ctypes.byref(nnz), None) == INVALID
    assert L.axb_bvh_query_nearest_rays(None, dummy, 4, 2, _policy(), dummy, ctypes.byref(ind),
                                        ctypes.byref(nnz), None) == INVALID
    assert L.axb_bruteforce_query_nearest_rays(None, dummy, 4, 2, _policy(), dummy,
                                               ctypes.byref(ind), ctypes.byref(nnz),
                                               None) == INVALID
    assert L.axb_bruteforce_query_nearest_rays(dummy, dummy, -1, 2, _policy(), dummy,
                                               ctypes.byref(ind), ctypes.byref(nnz),
                                               None) == INVALID


def test_energy_rejects_bad_arguments():
    L = ax.lib()
    dummy = P(16)
    assert L.axb_pair_energy(dummy, 10, dummy, dummy, 0.0, None, None, None) == INVALID
    assert "cutoff" in _last_error()
    assert L.axb_pair_energy(None, 10, dummy, dummy, 2.5, None, None, None) == INVALID
    assert L.axb_pair_energy(dummy, -1, dummy, dummy, 2.5, None, None, None) == INVALID


def test_size_and_bounds_of_null_handles():
    L = ax.lib()
    assert L.axb_bvh_size(None) == 0
    assert L.axb_bruteforce_size(None) == 0
    out = (ctypes.c_float * 6)()
    assert L.axb_bvh_bounds(None, ctypes.cast(out, P)) == INVALID
    assert L.axb_bruteforce_bounds(None, ctypes.cast(out, P)) == INVALID
    L.axb_bvh_destroy(None)  # no-ops
    L.axb_bruteforce_destroy(None)


def test_sorted_predicate_batches_below_2_30():
    """The predicate sort's look-back words hold 30-bit counts: a batch of
    2^30 or more predicates cannot be sorted by Morton code, and asking for
    it is an argument error (checked before any device work)."""
    L = ax.lib()
    ind = P()
    dist = P()
    nnz = I64()
    dummy = P(16)
    big = 1 << 30
    sort = ax.TraversalPolicy(sort_predicates=True)._c()
    assert L.axb_bvh_query_intersects(dummy, ax.PREDICATE_SPHERES, dummy, big, sort, dummy,
                                      ctypes.byref(ind), ctypes.byref(nnz), None) == INVALID
    assert "2^30" in _last_error()
    assert L.axb_bvh_query_intersects_radius(dummy, dummy, big, 0.5, sort, dummy,
                                             ctypes.byref(ind), ctypes.byref(nnz),
                                             None) == INVALID
    assert L.axb_bvh_query_nearest(dummy, dummy, big, 4, sort, dummy, ctypes.byref(ind),
                                   ctypes.byref(dist), ctypes.byref(nnz), None) == INVALID
    assert "2^30" in _last_error()
    assert L.axb_bvh_predicate_order(dummy, ctypes.c_int32(ax.PREDICATE_SPHERES), dummy,
                                     I64(big), dummy, None) == INVALID


def test_debug_limits_arguments():
    L = ax.lib()
    assert L.axb_debug_set_limits(-1, 0, 0) == INVALID
    assert L.axb_debug_set_limits(0, -1, 0) == INVALID
    assert L.axb_debug_set_limits(0, 0, 65) == INVALID
    assert L.axb_debug_set_limits(0, 0, -2) == INVALID
    assert L.axb_debug_set_limits(0, 0, 0) == ax.AXB_OK
