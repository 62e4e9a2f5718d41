"""GPU parity of the neighbour lists (findFullNeighborList,
findHalfNeighborList, expandHalfToFull; detail/ArborX_NeighborList.hpp,
detail/ArborX_ExpandHalfToFull.hpp) against the CPU oracle and the
reference's own known answers (test/tstNeighborList.cpp).  The reference
fills rows through atomics, so rows compare as multisets
(test/boost_ext/CompressedStorageComparison.hpp); the half list's pairs
(which of two points lists the other) follow the tree's leaf order, which is
bit-exact, so its rows must match too."""
import numpy as np
import pytest
import torch

import arborx_amd as ax
import oracle
from arborx_amd import synthetic
from tests.helpers import assert_crs_equal_as_sets

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def _r(r):
    return np.float32(np.inf) if r == "inf" else np.float32(r)


def test_neighbor_list_reference_cases(kat):
    nl = kat["neighbor_list"]
    for case in nl["degenerate"]:
        pts = np.array(case["points"], np.float32).reshape(-1, 3)
        for f in (ax.find_full_neighbor_list, ax.find_half_neighbor_list):
            o, i = f(dev(pts), _r(case["radius"]))
            assert list(host(o)) == case["offsets"] and host(i).size == 0
    pts = dev(np.array(nl["points"], np.float32))
    for case in nl["cases"]:
        r = _r(case["radius"])
        o, i = ax.find_full_neighbor_list(pts, r)
        assert_crs_equal_as_sets(host(o), host(i), case["offsets"], case["indices"])
        oh, ih = ax.find_half_neighbor_list(pts, r)
        o2, i2 = ax.expand_half_to_full(oh, ih)
        assert_crs_equal_as_sets(host(o2), host(i2), case["offsets"], case["indices"])


def _cloud(kind, n):
    if kind == "uniform":
        return synthetic.uniform_points(n, 21)
    if kind == "gaussian":
        return synthetic.gaussian_mixture_points(n, 22, n_centres=max(1, n // 2000))
    v = synthetic.uniform_points(n, 23)  # exact duplicates and a corner pile-up
    v[::5] = v[3]
    v[1::7] = np.float32(1.0)
    return v


@pytest.mark.parametrize("kind,n,nbrs", [("uniform", 100, 30), ("uniform", 20000, 16),
                                         ("gaussian", 30000, 20), ("dups", 5000, 12)])
@pytest.mark.parametrize("fmt", ["fp32", "q16"])
def test_neighbor_list_vs_oracle(kind, n, nbrs, fmt):
    pts = _cloud(kind, n)
    r = np.float32(synthetic.radius_for_neighbours(n, nbrs))
    of, if_ = oracle.neighbor_list(pts, r, True)
    oh, ih = oracle.neighbor_list(pts, r, False)
    bvh = ax.BoundingVolumeHierarchy(dev(pts), node_format=fmt)
    o, i = bvh.neighbor_list(r, full=True)
    assert_crs_equal_as_sets(host(o), host(i), of, if_)
    o, i = bvh.neighbor_list(r, full=False)
    assert_crs_equal_as_sets(host(o), host(i), oh, ih)
    o2, i2 = ax.expand_half_to_full(o, i)
    assert_crs_equal_as_sets(host(o2), host(i2), of, if_)
    # the free functions build their own (Q16) tree
    o, i = ax.find_full_neighbor_list(dev(pts), r)
    assert_crs_equal_as_sets(host(o), host(i), of, if_)


def test_neighbor_list_buffer_policies():
    n = 20000
    pts = _cloud("uniform", n)
    r = np.float32(synthetic.radius_for_neighbours(n, 24))
    of, if_ = oracle.neighbor_list(pts, r, True)
    bvh = ax.BoundingVolumeHierarchy(dev(pts), node_format="q16")
    ref = bvh.neighbor_list(r, full=True)
    # rows in traversal order: every policy gives the same arrays
    for b in (4, 16, 200, -200):
        o, i = bvh.neighbor_list(r, full=True, policy=ax.TraversalPolicy().setBufferSize(b))
        assert torch.equal(o, ref[0]) and torch.equal(i, ref[1])
    assert_crs_equal_as_sets(host(ref[0]), host(ref[1]), of, if_)
    with pytest.raises(ax.SearchException):
        bvh.neighbor_list(r, full=True, policy=ax.TraversalPolicy().setBufferSize(-2))
    o, i = bvh.neighbor_list(r, full=True)
    assert torch.equal(i, ref[1])


def test_neighbor_list_matches_radius_query_minus_self():
    """Full list = the radius query of the points themselves without the
    point's own leaf (exact duplicates stay)."""
    n = 50000
    pts = _cloud("dups", n)
    r = np.float32(synthetic.radius_for_neighbours(n, 16))
    bvh = ax.BoundingVolumeHierarchy(dev(pts), node_format="q16")
    qi, qo = bvh.query(ax.make_intersects(dev(pts), r))
    qo, qi = host(qo), host(qi)
    o, i = bvh.neighbor_list(r, full=True)
    o, i = host(o), host(i)
    np.testing.assert_array_equal(np.diff(o), np.diff(qo) - 1)
    for q in range(0, n, 97):
        row = qi[qo[q]:qo[q + 1]]
        k = np.flatnonzero(row == q)
        assert k.size == 1
        np.testing.assert_array_equal(i[o[q]:o[q + 1]], np.delete(row, k[0]))


def test_neighbor_list_rejects_bad_input():
    pts = dev(synthetic.uniform_points(100, 1))
    with pytest.raises(RuntimeError):
        ax.find_full_neighbor_list(pts, -1.0)
    boxes = ax.BoundingVolumeHierarchy(dev(synthetic.lognormal_boxes(100, 1)))
    with pytest.raises(RuntimeError):
        boxes.neighbor_list(0.1)
