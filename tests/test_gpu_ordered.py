"""GPU: ordered_intersects of spheres, boxes and points (OrderedSpatial<G>,
detail/ArborX_Predicates.hpp:104-155; TreeTraversal.hpp:338-489) over point
and box trees, both node formats, through the C ABI: every row lists all
values by (distance(geometry, value), index), equal to the oracle's brute
force; the reference's PerThread case (tstQueryTreeCallbackQueryPerThread.cpp:
81-110) as a fixture; degenerate trees and the hard buffer size."""
import numpy as np
import pytest
import torch

import arborx_amd as ax
import oracle

pytestmark = pytest.mark.gpu

KINDS = {"spheres": (ax.Spheres, oracle.ORDERED_SPHERES),
         "boxes": (ax.Boxes, oracle.ORDERED_BOXES),
         "points": (ax.Points, oracle.ORDERED_POINTS)}


def _values(geometry, n, rng):
    if geometry == 0:
        v = rng.uniform(0, 1, (n, 3)).astype(np.float32)
        v[7] = v[3]  # an exact duplicate: equal distances, ties by index
        return v
    lo = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    return np.concatenate([lo, lo + rng.uniform(0, 0.1, (n, 3)).astype(np.float32)], 1)


def _preds(kind, nq, rng):
    c = rng.uniform(-0.2, 1.2, (nq, 3)).astype(np.float32)
    if kind == "spheres":
        return np.concatenate([c, rng.uniform(0, 0.2, (nq, 1)).astype(np.float32)], 1)
    if kind == "boxes":
        return np.concatenate([c, c + rng.uniform(0, 0.2, (nq, 3)).astype(np.float32)], 1)
    return c


@pytest.mark.parametrize("fmt", ["fp32", "q16"])
@pytest.mark.parametrize("geometry", [0, 1])
@pytest.mark.parametrize("kind", ["spheres", "boxes", "points"])
def test_ordered_rows_equal_bruteforce(kind, geometry, fmt):
    rng = np.random.default_rng(7 + geometry * 3 + list(KINDS).index(kind))
    n, nq = 700, 257  # rows above 64 entries: the in-place heap sort
    v = _values(geometry, n, rng)
    p = _preds(kind, nq, rng)
    cls, okind = KINDS[kind]
    bvh = ax.BoundingVolumeHierarchy(torch.from_numpy(v).cuda(), node_format=fmt)
    idx, off = bvh.query(ax.ordered_intersects(cls(torch.from_numpy(p).cuda())))
    bo, bi = oracle.bruteforce_ordered(v, geometry, okind, p)
    np.testing.assert_array_equal(off.cpu().numpy(), bo)
    np.testing.assert_array_equal(idx.cpu().numpy(), bi)


@pytest.mark.parametrize("n", [1, 2, 5, 33, 64])
def test_ordered_rows_short_trees(n):
    """Rows of up to 64 values sort in registers; trees of 1 to 5 leaves have
    no stored internal node."""
    rng = np.random.default_rng(n)
    v = _values(0, max(n, 8), rng)[:n]
    p = _preds("boxes", 50, rng)
    for fmt in ("fp32", "q16"):
        bvh = ax.BoundingVolumeHierarchy(torch.from_numpy(v).cuda(), node_format=fmt)
        idx, off = bvh.query(ax.ordered_intersects(ax.Boxes(torch.from_numpy(p).cuda())))
        bo, bi = oracle.bruteforce_ordered(v, 0, oracle.ORDERED_BOXES, p)
        np.testing.assert_array_equal(off.cpu().numpy(), bo)
        np.testing.assert_array_equal(idx.cpu().numpy(), bi)


def test_per_thread_fixture_as_rows(kat):
    f = kat["per_thread"]
    pts = torch.tensor(f["points"], dtype=torch.float32).cuda()
    qb = torch.tensor(f["query_boxes"], dtype=torch.float32).cuda()
    for fmt in ("fp32", "q16"):
        bvh = ax.BoundingVolumeHierarchy(pts, node_format=fmt)
        idx, off = bvh.query(ax.ordered_intersects(ax.Boxes(qb)))
        o = off.cpu().numpy()
        assert list(o) == [10 * i for i in range(11)]
        assert [int(idx[a]) for a in o[:-1]] == f["first_value"]


def test_ordered_rows_degenerate_and_policy():
    q = torch.tensor([[0.5, 0.5, 0.5, 0.1]], dtype=torch.float32).cuda()
    empty = ax.BoundingVolumeHierarchy(torch.empty((0, 3)).cuda())
    idx, off = empty.query(ax.ordered_intersects(ax.Spheres(q)))
    assert idx.numel() == 0 and off.cpu().tolist() == [0, 0]
    pts = torch.rand((100, 3), generator=torch.Generator().manual_seed(3)).cuda()
    bvh = ax.BoundingVolumeHierarchy(pts)
    with pytest.raises(ax.SearchException):
        bvh.query(ax.ordered_intersects(ax.Spheres(q)), policy=ax.TraversalPolicy(buffer_size=-10))
    idx, off = bvh.query(ax.ordered_intersects(ax.Spheres(q)),
                         policy=ax.TraversalPolicy(buffer_size=-100))
    assert off.cpu().tolist() == [0, 100]
    # no predicates
    idx, off = bvh.query(ax.ordered_intersects(ax.Points(torch.empty((0, 3)).cuda())))
    assert idx.numel() == 0 and off.cpu().tolist() == [0]
