"""GPU parity of the brute-force index (ArborX::BruteForce,
src/spatial/ArborX_BruteForce.hpp) through its C ABI: spatial rows against
the oracle's brute force (rows in value order on both sides), nearest rows
and distances bit-exact under the (distance, index) rule, and the reference
suites the reference also runs over BruteForce (test/CMakeLists.txt:106-124
instantiates tstQueryTree{Basic,Callbacks,Degenerate,ManufacturedSolution}
for it): degenerate trees, the manufactured solution, collinear nearest."""
import math

import numpy as np
import pytest
import torch

import arborx_amd as ax
import oracle
from arborx_amd import synthetic
from tests.helpers import (assert_crs_equal_as_sets, manufactured_grid, manufactured_queries,
                           spheres_from_points)

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("geometry", ["points", "boxes"])
def test_spatial_rows_match_oracle(geometry):
    g = 0 if geometry == "points" else 1
    v = (synthetic.uniform_points(12_000, 31) if g == 0
         else synthetic.lognormal_boxes(12_000, 32, mu=math.log(0.01)))
    bf = ax.BruteForce(dev(v), geometry=geometry)
    # spheres with per-query radii, some zero
    q = spheres_from_points(synthetic.uniform_points(5_000, 33), 0.03)
    q[::7, 3] = 0.0
    o_ref, i_ref = oracle.bruteforce_spatial(v, g, 0, q)
    i, o = bf.query(ax.intersects(ax.Spheres(dev(q))))
    np.testing.assert_array_equal(host(o), o_ref)
    np.testing.assert_array_equal(host(i), i_ref)
    # boxes
    qb = synthetic.lognormal_boxes(5_000, 34, mu=math.log(0.03))
    o_ref, i_ref = oracle.bruteforce_spatial(v, g, 1, qb)
    i, o = bf.query(ax.intersects(ax.Boxes(dev(qb))))
    np.testing.assert_array_equal(host(o), o_ref)
    np.testing.assert_array_equal(host(i), i_ref)
    assert o_ref[-1] > 0


def test_radius_query_matches_tree_rows_as_sets():
    v = synthetic.uniform_points(20_000, 35)
    r = float(np.float32((48.0 / (4.0 * math.pi * 20_000)) ** (1.0 / 3.0)))
    bf = ax.BruteForce(dev(v))
    bvh = ax.BoundingVolumeHierarchy(dev(v), node_format="q16")
    i, o = bf.query(ax.make_intersects(dev(v), r))
    ti, to = bvh.query(ax.make_intersects(dev(v), r))
    np.testing.assert_array_equal(host(o), host(to))
    assert_crs_equal_as_sets(host(o), host(i), host(to), host(ti))
    o_ref, i_ref = oracle.bruteforce_spatial(v, 0, 0, spheres_from_points(v, np.float32(r)))
    np.testing.assert_array_equal(host(i), i_ref)


@pytest.mark.parametrize("k", [1, 3, 16, 33, 64, 100])
@pytest.mark.parametrize("geometry", ["points", "boxes"])
def test_nearest_bit_exact(k, geometry):
    g = 0 if geometry == "points" else 1
    v = (synthetic.uniform_points(8_000, 36) if g == 0
         else synthetic.lognormal_boxes(8_000, 37, mu=math.log(0.01)))
    q = synthetic.uniform_points(3_000, 38)
    i_ref, d_ref = oracle.bruteforce_knn(v, g, q, k)
    bf = ax.BruteForce(dev(v), geometry=geometry)
    i, o, d = bf.query(ax.make_nearest(dev(q), k), return_distances=True)
    assert host(o).tolist() == [j * min(k, 8_000) for j in range(3_001)]
    np.testing.assert_array_equal(host(i), i_ref)
    np.testing.assert_array_equal(host(d).view(np.uint32), d_ref.view(np.uint32))
    # without distances (k > 64 takes a temporary for the heap)
    i2, _ = bf.query(ax.make_nearest(dev(q), k))
    np.testing.assert_array_equal(host(i2), i_ref)


def test_nearest_ties_match_tree():
    """Integer grid: many exact distance ties, and queries at half-integers
    where d2 of different points round to one distance."""
    gr = np.arange(14, dtype=np.float32)
    v = np.stack(np.meshgrid(gr, gr, gr, indexing="ij"), -1).reshape(-1, 3).copy()
    q = np.floor(synthetic.uniform_points(2000, 43) * 14).astype(np.float32) + np.float32(0.5)
    q[::2] = np.floor(q[::2])
    q[1::3] += np.float32(1e-3)
    bf = ax.BruteForce(dev(v))
    bvh = ax.BoundingVolumeHierarchy(dev(v))
    for k in (7, 27, 64, 100):
        i_ref, d_ref = oracle.bruteforce_knn(v, 0, q, k)
        i, o, d = bf.query(ax.make_nearest(dev(q), k), return_distances=True)
        np.testing.assert_array_equal(host(i), i_ref)
        np.testing.assert_array_equal(host(d), d_ref)
        ti, to = bvh.query(ax.make_nearest(dev(q), k))
        np.testing.assert_array_equal(host(i), host(ti))


def test_bounds_size_and_empty():
    v = synthetic.lognormal_boxes(1000, 39)
    bf = ax.BruteForce(dev(v), geometry="boxes")
    bvh = ax.BoundingVolumeHierarchy(dev(v), geometry="boxes")
    assert bf.size() == 1000 and not bf.empty()
    assert bf.bounds() == bvh.bounds()
    e = ax.BruteForce()
    assert e.empty() and e.size() == 0
    FM = float(np.finfo(np.float32).max)
    assert e.bounds() == [FM, FM, FM, -FM, -FM, -FM]
    q = dev(synthetic.uniform_points(10, 40))
    i, o = e.query(ax.make_nearest(q, 3))
    assert host(o).tolist() == [0] * 11 and i.numel() == 0
    i, o = e.query(ax.make_intersects(q, 0.5))
    assert host(o).tolist() == [0] * 11 and i.numel() == 0
    one = ax.BruteForce(dev(np.array([[0.5, 0.5, 0.5]], np.float32)))
    i, o, d = one.query(ax.make_nearest(q, 5), return_distances=True)
    assert host(i).tolist() == [0] * 10
    i, o = one.query(ax.make_nearest(q, 0))
    assert host(o).tolist() == [0] * 11


def test_degenerate_cases(kat):
    for c in kat["degenerate"]["cases"]:
        tree = np.array(c["tree"], np.float32).reshape(-1, 6)
        bf = ax.BruteForce(dev(tree), geometry="boxes")
        if c["kind"] == "nearest":
            preds = np.array(c["preds"], np.float64).reshape(-1, 4)
            off, idx = [0], []
            for p in preds:
                i, o = bf.query(ax.make_nearest(dev(p[None, :3].astype(np.float32)), int(p[3])))
                idx += host(i).tolist()
                off.append(len(idx))
        else:
            if c["kind"] == "spheres":
                preds = np.array(c["preds"], np.float32).reshape(-1, 4)
                i, o = bf.query(ax.intersects(ax.Spheres(dev(preds))))
            else:
                preds = np.array(c["preds"], np.float32).reshape(-1, 6)
                i, o = bf.query(ax.intersects(ax.Boxes(dev(preds))))
            off, idx = host(o), host(i)
        assert_crs_equal_as_sets(off, idx, c["offsets"], c["indices"])


def test_manufactured_solution():
    boxes, *_ = manufactured_grid()
    bf = ax.BruteForce(dev(boxes), geometry="boxes")
    for name, q, off, ind in manufactured_queries():
        i, o = bf.query(ax.intersects(ax.Boxes(dev(q))))
        assert_crs_equal_as_sets(host(o), host(i), off, ind)


def test_nearest_collinear(kat):
    a = kat["nearest_collinear"]
    bf = ax.BruteForce(dev(np.array(a["points"], np.float32)))
    i, o, d = bf.query(ax.make_nearest(dev(np.array([a["query"]], np.float32)), a["k"]),
                       return_distances=True)
    assert host(i).tolist() == a["indices"]
    np.testing.assert_array_equal(host(d), np.array(a["distances"], np.float32))


def test_buffer_policy():
    v = synthetic.uniform_points(10_000, 41)
    r = 0.08
    o_ref, i_ref = oracle.bruteforce_spatial(v, 0, 0, spheres_from_points(v, np.float32(r)))
    bf = ax.BruteForce(dev(v))
    pts = dev(v)
    for b in (64, 8, 0):
        i, o = bf.query(ax.make_intersects(pts, r), ax.TraversalPolicy().setBufferSize(b))
        np.testing.assert_array_equal(host(o), o_ref)
        np.testing.assert_array_equal(host(i), i_ref)
    with pytest.raises(ax.SearchException):
        bf.query(ax.make_intersects(pts, r), ax.TraversalPolicy().setBufferSize(-1))
    i, o = bf.query(ax.make_intersects(pts, r))
    np.testing.assert_array_equal(host(i), i_ref)
