"""The oracle's ordered (best-first) traversal for spheres, boxes and points
(Experimental::ordered_intersects over any tree, detail/ArborX_Predicates.hpp:
104-126, TreeTraversal.hpp:338-489) and the distances it orders by
(geometry/algorithms/ArborX_Distance.hpp:52-208), pinned by closed forms, the
reference's PerThread test and brute force."""
import math

import numpy as np
import pytest

import oracle

RAYS, SPHERES, BOXES, POINTS = (oracle.ORDERED_RAYS, oracle.ORDERED_SPHERES,
                                oracle.ORDERED_BOXES, oracle.ORDERED_POINTS)
f32 = np.float32


def d(kind, pred, geometry, value):
    return oracle.ordered_distance(kind, pred, geometry, value)


def test_box_box_distance_closed_forms():
    q = [0, 0, 0, 1, 1, 1]
    # apart along x only: the gap, on either side
    assert d(BOXES, q, 1, [3, 0, 0, 4, 1, 1]) == 2.0
    assert d(BOXES, q, 1, [-4, 0, 0, -2.5, 1, 1]) == 2.5
    # apart along x and y: sqrt(gx^2 + gy^2); along all three: the diagonal
    assert d(BOXES, q, 1, [4, 5, 0, 6, 6, 1]) == 5.0
    assert d(BOXES, q, 1, [2, 2, 2, 3, 3, 3]) == f32(math.sqrt(3.0))
    # overlapping or touching on every axis: 0; overlapping on two axes: the third gap
    assert d(BOXES, q, 1, [0.5, 0.5, 0.5, 2, 2, 2]) == 0.0
    assert d(BOXES, q, 1, [1, 1, 1, 2, 2, 2]) == 0.0
    assert d(BOXES, q, 1, [0.2, -3, 0.2, 0.3, -1, 0.3]) == 1.0


def test_sphere_distances_closed_forms():
    s = [0, 0, 0, 1]
    # to a box: distance of the centre minus r, never below 0
    assert d(SPHERES, s, 1, [4, 0, 0, 5, 1, 1]) == 3.0
    assert d(SPHERES, s, 1, [0.5, 0, 0, 5, 1, 1]) == 0.0
    assert d(SPHERES, s, 1, [-1, -1, -1, 1, 1, 1]) == 0.0
    # to a point (reverse dispatch: distance(point, sphere))
    assert d(SPHERES, s, 0, [3, 4, 0]) == 4.0
    assert d(SPHERES, s, 0, [0.5, 0, 0]) == 0.0
    assert d(SPHERES, [1, 1, 1, 0.25], 0, [1, 1, 3]) == 1.75


def test_point_distances_closed_forms():
    p = [0, 0, 0]
    assert d(POINTS, p, 0, [3, 4, 12]) == 13.0
    assert d(POINTS, p, 1, [3, 4, 12, 5, 5, 13]) == 13.0  # nearest corner
    assert d(POINTS, p, 1, [-1, 2, -1, 1, 3, 1]) == 2.0  # face
    assert d(POINTS, p, 1, [-1, -1, -1, 1, 1, 1]) == 0.0  # inside
    # a box predicate against a point value: distance(point, box)
    assert d(BOXES, [3, 4, 12, 5, 5, 13], 0, p) == 13.0
    assert d(BOXES, [-1, -1, -1, 1, 1, 1], 0, p) == 0.0


def test_point_values_equal_degenerate_boxes():
    """distance(geometry, point) equals distance(geometry, [p, p]) for the
    three predicate kinds (what lets the library treat a point leaf as a
    degenerate box)."""
    rng = np.random.default_rng(11)
    for _ in range(2000):
        p = rng.uniform(-2, 2, 3).astype(np.float32)
        pb = np.concatenate([p, p])
        c = rng.uniform(-2, 2, 3).astype(np.float32)
        lo = rng.uniform(-2, 2, 3).astype(np.float32)
        hi = lo + rng.uniform(0, 1, 3).astype(np.float32)
        for kind, q in ((SPHERES, np.append(c, f32(rng.uniform(0, 1)))),
                        (BOXES, np.concatenate([lo, hi])), (POINTS, c)):
            a, b = d(kind, q, 0, p), d(kind, q, 1, pb)
            assert np.float32(a).view(np.uint32) == np.float32(b).view(np.uint32), (kind, q, p)


def test_per_thread_ordered_intersects_fixture(kat):
    """tstQueryTreeCallbackQueryPerThread.cpp:81-110: the first value the
    ordered walk reports for box i is point i; :36-79: intersects(box i)
    matches point i only."""
    f = kat["per_thread"]
    pts = np.array(f["points"], np.float32)
    qb = np.array(f["query_boxes"], np.float32)
    t = oracle.Tree(pts)
    off, idx, dist = t.query_ordered(qb, BOXES)
    assert np.all(np.diff(off) == len(pts))  # every point, in order of distance
    assert [int(idx[o]) for o in off[:-1]] == f["first_value"]
    assert np.all(dist[off[:-1]] == 0.0)
    o2, i2 = t.query_spatial(1, qb)
    assert list(o2) == list(range(11)) and list(i2) == f["first_value"]


def _sorted_by_index_within_ties(off, idx, dist):
    out = idx.copy()
    for a, b in zip(off[:-1], off[1:]):
        order = np.lexsort((idx[a:b], dist[a:b]))
        out[a:b] = idx[a:b][order]
    return out


@pytest.mark.parametrize("geometry", [0, 1])
@pytest.mark.parametrize("kind", [SPHERES, BOXES, POINTS])
def test_ordered_traversal_equals_bruteforce(geometry, kind):
    rng = np.random.default_rng(100 + 10 * geometry + kind)
    n, nq = 40, 300
    if geometry == 0:
        values = rng.uniform(0, 1, (n, 3)).astype(np.float32)
        values[5] = values[6]  # an exact duplicate
    else:
        lo = rng.uniform(0, 1, (n, 3)).astype(np.float32)
        values = np.concatenate([lo, lo + rng.uniform(0, 0.2, (n, 3)).astype(np.float32)], 1)
    c = rng.uniform(-0.2, 1.2, (nq, 3)).astype(np.float32)
    if kind == SPHERES:
        preds = np.concatenate([c, rng.uniform(0, 0.3, (nq, 1)).astype(np.float32)], 1)
    elif kind == BOXES:
        preds = np.concatenate([c, c + rng.uniform(0, 0.3, (nq, 3)).astype(np.float32)], 1)
    else:
        preds = c
    t = oracle.Tree(values, geometry)
    off, idx, dist = t.query_ordered(preds, kind)
    bo, bi = oracle.bruteforce_ordered(values, geometry, kind, preds)
    np.testing.assert_array_equal(off, bo)
    assert np.all(np.diff(off) == n)  # every value is at a finite distance
    for a, b in zip(off[:-1], off[1:]):
        assert np.all(np.diff(dist[a:b]) >= 0)
    np.testing.assert_array_equal(_sorted_by_index_within_ties(off, idx, dist), bi)
    # the rows' distances are those of the brute-force formula
    for q in range(0, nq, 37):
        for j in range(off[q], off[q + 1]):
            assert dist[j] == d(kind, preds[q], geometry, values[idx[j]])
