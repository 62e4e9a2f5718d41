"""GPU parity of the ray predicates through the C ABI: Intersects<Ray>,
OrderedSpatial<Ray> (the best-first traversal, TreeTraversal.hpp:338-489)
and Nearest<Ray> over box trees of both node formats and over the brute-force
index, against the oracle: the reference's ray tests
(test/tstQueryTreeRay.cpp) as known answers, brute-force rows in
(distance, index) order, the oracle's reference-form rope traversal for the
unordered rows (same order within a row), and the ray construction
bit-exact."""
import math

import numpy as np
import pytest
import torch

import arborx_amd as ax
import oracle
from arborx_amd import synthetic
from tests.helpers import assert_crs_equal_as_sets, ray_inputs

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def test_make_rays_bit_exact():
    o, d = ray_inputs(20_000, 5)
    d[::11] *= np.float32(1e-20)  # tiny but non-zero directions
    want = oracle.make_rays(o, d)
    r = ax.Rays(dev(o), dev(d))
    np.testing.assert_array_equal(host(r.data).view(np.uint32), want.view(np.uint32))
    with pytest.raises(ValueError):
        ax.Rays(dev(np.zeros((2, 3), np.float32)), dev(np.array([[1, 0, 0], [0, 0, 0]],
                                                                np.float32)))


@pytest.mark.parametrize("fmt", ["fp32", "q16"])
def test_ray_kat(kat, fmt):
    for c in kat["rays"]["cases"]:
        boxes = np.array(c["boxes"], np.float32).reshape(-1, 6)
        bvh = ax.BoundingVolumeHierarchy(dev(boxes) if len(boxes) else None, geometry="boxes",
                                         node_format=fmt)
        rays = ax.Rays(dev(np.array(c["origins"], np.float32)),
                       dev(np.array(c["directions"], np.float32)))
        if c["kind"] == "nearest":
            i, o = bvh.query(ax.make_nearest(rays, c["k"]))
        elif c["kind"] == "ordered":
            i, o = bvh.query(ax.make_ordered_intersects(rays))
        else:
            i, o = bvh.query(ax.intersects(rays))
            assert_crs_equal_as_sets(host(o), host(i), c["offsets"], c["indices"])
            continue
        assert host(o).tolist() == c["offsets"], c["line"]
        assert host(i).tolist() == c["indices"], c["line"]


@pytest.mark.parametrize("fmt", ["fp32", "q16"])
def test_ray_rows_match_oracle(fmt):
    boxes = synthetic.lognormal_boxes(40_000, 71, mu=math.log(0.01))
    o, d = ray_inputs(4_000, 72, boxes)
    rays_h = oracle.make_rays(o, d)
    rays = ax.Rays.from_normalized(dev(rays_h))
    bvh = ax.BoundingVolumeHierarchy(dev(boxes), geometry="boxes", node_format=fmt)
    t = oracle.Tree(boxes, geometry=1)
    # unordered: the reference's rope traversal order within each row
    so, si = t.query_spatial(3, rays_h)
    i, off = bvh.query(ax.intersects(rays))
    np.testing.assert_array_equal(host(off), so)
    np.testing.assert_array_equal(host(i), si)
    # ordered: (distance, index)
    bo, bi = oracle.bruteforce_rays_ordered(boxes, rays_h)
    i, off = bvh.query(ax.ordered_intersects(rays))
    np.testing.assert_array_equal(host(off), bo)
    np.testing.assert_array_equal(host(i), bi)
    assert bo[-1] > 4_000
    # nearest(ray, k): the first k of each ordered row
    for k in (1, 3, 10):
        ko, ki = oracle.bruteforce_rays_ordered(boxes, rays_h, k)
        i, off = bvh.query(ax.nearest(rays, k))
        np.testing.assert_array_equal(host(off), ko)
        np.testing.assert_array_equal(host(i), ki)
    i, off = bvh.query(ax.nearest(rays, 0))
    assert i.numel() == 0 and host(off).tolist() == [0] * (len(rays_h) + 1)


@pytest.mark.parametrize("fmt", ["fp32", "q16"])
def test_long_ray_rows(fmt):
    """Large overlapping boxes: rows of hundreds of hits (ordered rows past
    the 64 sorted in registers take the in-place heap sort; nearest(ray, k)
    walks long best-first frontiers)."""
    boxes = synthetic.lognormal_boxes(30_000, 80, mu=math.log(0.08))
    o, d = ray_inputs(1_000, 81, boxes)
    rays_h = oracle.make_rays(o, d)
    rays = ax.Rays.from_normalized(dev(rays_h))
    bvh = ax.BoundingVolumeHierarchy(dev(boxes), geometry="boxes", node_format=fmt)
    bo, bi = oracle.bruteforce_rays_ordered(boxes, rays_h)
    assert np.percentile(np.diff(bo), 50) > 64
    i, off = bvh.query(ax.ordered_intersects(rays))
    np.testing.assert_array_equal(host(off), bo)
    np.testing.assert_array_equal(host(i), bi)
    ko, ki = oracle.bruteforce_rays_ordered(boxes, rays_h, 100)
    i, off = bvh.query(ax.nearest(rays, 100))
    np.testing.assert_array_equal(host(off), ko)
    np.testing.assert_array_equal(host(i), ki)


@pytest.mark.parametrize("fmt", ["fp32", "q16"])
def test_nearest_rays_past_the_heap(fmt):
    """Boxes up to 0.4 wide: a best-first walk of many rays needs more than
    its 64 pending entries (the CPU test of the walk counts them); the query
    then takes the first k of the whole ordered rows, with the same rows."""
    g = np.random.default_rng(82)
    lo = g.uniform(size=(30_000, 3)).astype(np.float32)
    e = (0.004 + 0.4 * g.uniform(size=(30_000, 1)) * g.uniform(size=(30_000, 1)))
    boxes = np.concatenate([lo, lo + e.astype(np.float32)], axis=1)
    o, d = ray_inputs(600, 83, boxes)
    rays_h = oracle.make_rays(o, d)
    rays = ax.Rays.from_normalized(dev(rays_h))
    bvh = ax.BoundingVolumeHierarchy(dev(boxes), geometry="boxes", node_format=fmt)
    for k in (1, 50, 400):
        ko, ki = oracle.bruteforce_rays_ordered(boxes, rays_h, k)
        i, off = bvh.query(ax.nearest(rays, k))
        np.testing.assert_array_equal(host(off), ko)
        np.testing.assert_array_equal(host(i), ki)


@pytest.mark.parametrize("buffer_size", [0, 3, 64, -64])
@pytest.mark.parametrize("sort", [True, False])
def test_ordered_policies(buffer_size, sort):
    boxes = synthetic.lognormal_boxes(20_000, 73, mu=math.log(0.02))
    o, d = ray_inputs(2_000, 74, boxes)
    rays_h = oracle.make_rays(o, d)
    rays = ax.Rays.from_normalized(dev(rays_h))
    bvh = ax.BoundingVolumeHierarchy(dev(boxes), geometry="boxes", node_format="q16")
    bo, bi = oracle.bruteforce_rays_ordered(boxes, rays_h)
    pol = ax.TraversalPolicy().setBufferSize(buffer_size).setPredicateSorting(sort)
    longest = int(np.max(np.diff(bo)))
    if buffer_size < 0 and longest > -buffer_size:
        with pytest.raises(ax.SearchException):
            bvh.query(ax.ordered_intersects(rays), pol)
        return
    i, off = bvh.query(ax.ordered_intersects(rays), pol)
    np.testing.assert_array_equal(host(off), bo)
    np.testing.assert_array_equal(host(i), bi)
    ko, ki = oracle.bruteforce_rays_ordered(boxes, rays_h, 5)
    i, off = bvh.query(ax.nearest(rays, 5), pol)
    np.testing.assert_array_equal(host(i), ki)


def test_rays_at_scale_properties():
    """One million boxes, 200k rays: ordered rows are the unordered rows
    sorted by (distance, index) -- checked as sets plus the oracle's rows for
    a sample of rays -- and nearest(ray, 4) rows are their prefixes."""
    boxes = synthetic.lognormal_boxes(1_000_000, 75)
    o, d = ray_inputs(200_000, 76, boxes)
    rays_h = oracle.make_rays(o, d)
    rays = ax.Rays.from_normalized(dev(rays_h))
    bvh = ax.BoundingVolumeHierarchy(dev(boxes), geometry="boxes", node_format="q16")
    ui, uo = bvh.query(ax.intersects(rays))
    oi, oo = bvh.query(ax.ordered_intersects(rays))
    uo, ui, oo, oi = host(uo), host(ui), host(oo), host(oi)
    np.testing.assert_array_equal(uo, oo)
    assert uo[-1] > 100_000
    row = np.repeat(np.arange(len(uo) - 1), np.diff(uo))
    a = np.lexsort((ui, row))
    b = np.lexsort((oi, row))
    np.testing.assert_array_equal(ui[a], oi[b])
    sample = np.arange(0, len(rays_h), 997)
    so, si = oracle.bruteforce_rays_ordered(boxes, rays_h[sample])
    got = np.concatenate([oi[oo[j]:oo[j + 1]] for j in sample])
    np.testing.assert_array_equal(got, si)
    ni, no = bvh.query(ax.nearest(rays, 4))
    ni, no = host(ni), host(no)
    np.testing.assert_array_equal(np.diff(no), np.minimum(np.diff(oo), 4))
    want = np.concatenate([oi[oo[j]:oo[j] + 4][:no[j + 1] - no[j]] for j in range(0, len(oo) - 1,
                                                                                  101)])
    got = np.concatenate([ni[no[j]:no[j + 1]] for j in range(0, len(no) - 1, 101)])
    np.testing.assert_array_equal(got, want)


def test_bruteforce_rays():
    boxes = synthetic.lognormal_boxes(10_000, 77, mu=math.log(0.02))
    o, d = ray_inputs(3_000, 78, boxes)
    rays_h = oracle.make_rays(o, d)
    rays = ax.Rays.from_normalized(dev(rays_h))
    bf = ax.BruteForce(dev(boxes), geometry="boxes")
    bo, bi = oracle.bruteforce_spatial(boxes, 1, 3, rays_h)
    i, off = bf.query(ax.intersects(rays))
    np.testing.assert_array_equal(host(off), bo)
    np.testing.assert_array_equal(host(i), bi)
    with pytest.raises(RuntimeError):
        bf.query(ax.ordered_intersects(rays))
    # nearest(ray, k): the first k boxes by (distance, index), rows may be short
    for k, b in ((1, 0), (5, 0), (12, 0), (5, 3), (12, -16)):
        ko, ki = oracle.bruteforce_rays_ordered(boxes, rays_h, k)
        i, off = bf.query(ax.nearest(rays, k), ax.TraversalPolicy().setBufferSize(b))
        np.testing.assert_array_equal(host(off), ko)
        np.testing.assert_array_equal(host(i), ki)
    assert np.max(np.diff(ko)) == 12 and np.min(np.diff(ko)) < 12
    # long rows (a heap of 100 over boxes up to 0.1 wide) against the tree
    big = synthetic.lognormal_boxes(4_000, 84, mu=math.log(0.1))
    bfb = ax.BruteForce(dev(big), geometry="boxes")
    ko, ki = oracle.bruteforce_rays_ordered(big, rays_h[:500], 100)
    sub = ax.Rays.from_normalized(dev(rays_h[:500]))
    i, off = bfb.query(ax.nearest(sub, 100))
    np.testing.assert_array_equal(host(off), ko)
    np.testing.assert_array_equal(host(i), ki)
    assert np.max(np.diff(ko)) == 100


def test_rays_need_boxes():
    pts = synthetic.uniform_points(1000, 79)
    rays = ax.Rays(dev(pts[:10]), dev(pts[10:20] - np.float32(0.5)))
    bvh = ax.BoundingVolumeHierarchy(dev(pts))
    for p in (ax.intersects(rays), ax.ordered_intersects(rays), ax.nearest(rays, 2)):
        with pytest.raises(RuntimeError, match="boxes"):
            bvh.query(p)
    for p in (ax.intersects(rays), ax.nearest(rays, 2)):
        with pytest.raises(RuntimeError):
            ax.BruteForce(dev(pts)).query(p)


def test_ray_edge_cases():
    """No rays, k = 0, empty indexes, a single box (TreeTraversal.hpp:48-59,
    143-178: degenerate trees and k < 1)."""
    none = ax.Rays(dev(np.zeros((0, 3), np.float32)), dev(np.zeros((0, 3), np.float32)))
    assert none.data.shape == (0, 6)
    one = dev(np.array([[0, 0, 0, 1, 1, 1]], np.float32))
    rays = ax.Rays(dev(np.array([[0.5, 0.5, -1.0], [2.0, 2.0, 2.0]], np.float32)),
                   dev(np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0]], np.float32)))
    for index in (ax.BoundingVolumeHierarchy(one, geometry="boxes", node_format="q16"),
                  ax.BruteForce(one, geometry="boxes")):
        i, off = index.query(ax.intersects(none))
        assert host(off).tolist() == [0] and i.numel() == 0
        i, off = index.query(ax.nearest(rays, 0))
        assert host(off).tolist() == [0, 0, 0] and i.numel() == 0
        i, off = index.query(ax.nearest(rays, 3))
        assert host(off).tolist() == [0, 1, 1] and host(i).tolist() == [0]
        i, off = index.query(ax.intersects(rays))
        assert host(off).tolist() == [0, 1, 1] and host(i).tolist() == [0]
    bvh = ax.BoundingVolumeHierarchy(one, geometry="boxes")
    i, off = bvh.query(ax.ordered_intersects(rays))
    assert host(off).tolist() == [0, 1, 1] and host(i).tolist() == [0]
    for empty in (ax.BoundingVolumeHierarchy(), ax.BruteForce()):
        for p in (ax.intersects(rays), ax.nearest(rays, 2)):
            i, off = empty.query(p)
            assert host(off).tolist() == [0, 0, 0] and i.numel() == 0
