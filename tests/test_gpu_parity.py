"""GPU parity: the HIP library (through its C ABI) against the CPU oracle on
the same seeded inputs.  Morton codes, permutation, topology and node arrays
must be bit-exact; result rows must be identical (same traversal order as the
reference's rope walk, so even the order within a row matches); kNN rows and
distances bit-exact under the (distance, index) rule."""
import math

import numpy as np
import pytest
import torch

import arborx_amd as ax
import oracle
from arborx_amd import synthetic
from tests.helpers import (assert_crs_equal_as_sets, diagonal_boxes, manufactured_grid,
                           manufactured_queries, spheres_from_points)

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def dataset(kind, n):
    if kind == "uniform":
        return synthetic.uniform_points(n, 2), 0
    if kind == "gaussian":
        return synthetic.gaussian_mixture_points(n, 3, n_centres=max(1, n // 10000)), 0
    if kind == "duplicates":
        v = synthetic.uniform_points(n, 9)
        v[::4] = v[7]
        v[1::5] = np.float32(1.0)
        return v, 0
    if kind == "boxes":
        return synthetic.lognormal_boxes(n, 4), 1
    raise ValueError(kind)


_cache = {}


def oracle_tree(kind, n):
    key = (kind, n)
    if key not in _cache:
        v, g = dataset(kind, n)
        _cache[key] = (v, g, oracle.Tree(v, g))
    return _cache[key]


# ------------------------------------------------------------ build
@pytest.mark.parametrize("kind,n", [("uniform", 100_000), ("gaussian", 100_000),
                                    ("duplicates", 50_000), ("boxes", 100_000),
                                    ("uniform", 2), ("uniform", 3), ("uniform", 1000)])
def test_build_topology_bit_exact(kind, n):
    v, g, t = oracle_tree(kind, n)
    bvh = ax.BoundingVolumeHierarchy(dev(v), geometry="points" if g == 0 else "boxes",
                                     keep_topology=True)
    assert bvh.size() == n
    np.testing.assert_array_equal(np.array(bvh.bounds(), np.float32), t.bounds())
    e = bvh.export_topology()
    r = t.export()
    np.testing.assert_array_equal(host(e["keys"]).view(np.uint64), r["keys"])
    np.testing.assert_array_equal(host(e["perm"]).view(np.uint32), r["perm"])
    np.testing.assert_array_equal(host(e["leaf_rope"]), r["leaf_rope"])
    np.testing.assert_array_equal(host(e["left_child"]), r["left_child"])
    np.testing.assert_array_equal(host(e["rope"]), r["rope"])
    np.testing.assert_array_equal(host(e["boxes"]).view(np.uint32), r["boxes"].view(np.uint32))


@pytest.mark.parametrize("node_format", ["fp32", "q16"])
@pytest.mark.parametrize("kind,n", [("uniform", 100_000), ("gaussian", 100_000),
                                    ("boxes", 50_000), ("uniform", 1), ("uniform", 2)])
def test_preorder_nodes_bit_exact(node_format, kind, n):
    v, g, t = oracle_tree(kind, n)
    bvh = ax.BoundingVolumeHierarchy(dev(v), geometry="points" if g == 0 else "boxes",
                                     node_format=node_format)
    ref = t.linearize(0 if node_format == "fp32" else 1, flat_max=ax.FLAT_SUBTREE_MAX)
    got = host(bvh.slots()).view(np.uint32)
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got, ref)


def test_trim_memory_between_queries():
    """Cached workspaces and pool memory handed back between queries: the
    next queries allocate afresh and give the same rows."""
    pts = dev(synthetic.uniform_points(50_000, 31))
    r = float(np.float32(synthetic.radius_for_neighbours(50_000, 16)))
    bvh = ax.BoundingVolumeHierarchy(pts, node_format="q16")
    i1, o1 = bvh.query(ax.make_intersects(pts, r))
    k1, _ = bvh.query(ax.make_nearest(pts, 8))
    torch.cuda.synchronize()
    assert ax.lib().axb_trim_memory() == 0
    i2, o2 = bvh.query(ax.make_intersects(pts, r))
    k2, _ = bvh.query(ax.make_nearest(pts, 8))
    assert torch.equal(o1, o2) and torch.equal(i1, i2) and torch.equal(k1, k2)


def test_empty_tree():
    bvh = ax.BoundingVolumeHierarchy(torch.empty((0, 3), device="cuda"))
    assert bvh.empty() and bvh.size() == 0
    fm = float(np.finfo(np.float32).max)
    assert bvh.bounds() == [fm, fm, fm, -fm, -fm, -fm]
    idx, off = bvh.query(ax.make_intersects(torch.zeros((2, 3)), 1.0))
    assert host(off).tolist() == [0, 0, 0] and idx.numel() == 0
    idx, off = bvh.query(ax.make_nearest(torch.zeros((2, 3)), 3))
    assert host(off).tolist() == [0, 0, 0] and idx.numel() == 0


# ------------------------------------------------------------ spatial
@pytest.mark.parametrize("node_format", ["fp32", "q16"])
@pytest.mark.parametrize("kind", ["uniform", "gaussian", "duplicates"])
@pytest.mark.parametrize("sort_predicates", [True, False])
def test_radius_query_identical_rows(node_format, kind, sort_predicates):
    n = 100_000 if kind != "duplicates" else 50_000
    v, g, t = oracle_tree(kind, n)
    r = np.float32((48.0 / (4.0 * math.pi * n)) ** (1.0 / 3.0))
    o_ref, i_ref = t.query_spatial(0, spheres_from_points(v, r))
    bvh = ax.BoundingVolumeHierarchy(dev(v), node_format=node_format)
    pol = ax.TraversalPolicy().setPredicateSorting(sort_predicates)
    idx, off = bvh.query(ax.make_intersects(dev(v), float(r)), pol)
    np.testing.assert_array_equal(host(off), o_ref)
    np.testing.assert_array_equal(host(idx), i_ref)


@pytest.mark.parametrize("node_format", ["fp32", "q16"])
def test_sphere_query_per_query_radius(node_format):
    v, g, t = oracle_tree("uniform", 100_000)
    rng = np.random.default_rng(5)
    q = np.concatenate([synthetic.uniform_points(20_000, 77) * 1.2 - 0.1,
                        rng.uniform(0.0, 0.05, (20_000, 1)).astype(np.float32)], 1)
    q[::97, 3] = 0.0  # zero radius: only exact hits
    q[5, :3] = v[123]
    q[5, 3] = 0.0
    o_ref, i_ref = t.query_spatial(0, q)
    bvh = ax.BoundingVolumeHierarchy(dev(v), node_format=node_format)
    idx, off = bvh.query(ax.intersects(ax.Spheres(dev(q))))
    np.testing.assert_array_equal(host(off), o_ref)
    np.testing.assert_array_equal(host(idx), i_ref)
    assert 123 in host(idx)[host(off)[5]:host(off)[6]]


@pytest.mark.parametrize("node_format", ["fp32", "q16"])
@pytest.mark.parametrize("geometry", ["boxes", "points"])
def test_box_query_identical_rows(node_format, geometry):
    if geometry == "boxes":
        v, g, t = oracle_tree("boxes", 100_000)
        q = synthetic.lognormal_boxes(100_000, 5)
    else:
        v, g, t = oracle_tree("uniform", 100_000)
        q = synthetic.lognormal_boxes(50_000, 6, mu=math.log(0.02))
    o_ref, i_ref = t.query_spatial(1, q)
    bvh = ax.BoundingVolumeHierarchy(dev(v), geometry=geometry, node_format=node_format)
    idx, off = bvh.query(ax.intersects(ax.Boxes(dev(q))))
    np.testing.assert_array_equal(host(off), o_ref)
    np.testing.assert_array_equal(host(idx), i_ref)
    assert o_ref[-1] > 0


def test_sphere_query_on_boxes():
    v, g, t = oracle_tree("boxes", 100_000)
    q = spheres_from_points(synthetic.uniform_points(30_000, 8), 0.01)
    o_ref, i_ref = t.query_spatial(0, q)
    for fmt in ("fp32", "q16"):
        bvh = ax.BoundingVolumeHierarchy(dev(v), geometry="boxes", node_format=fmt)
        idx, off = bvh.query(ax.intersects(ax.Spheres(dev(q))))
        np.testing.assert_array_equal(host(off), o_ref)
        np.testing.assert_array_equal(host(idx), i_ref)


# ------------------------------------------------------------ nearest
@pytest.mark.parametrize("node_format", ["fp32", "q16"])
@pytest.mark.parametrize("k", [1, 10, 16, 32])
@pytest.mark.parametrize("kind", ["uniform", "gaussian", "duplicates"])
def test_knn_bit_exact(node_format, k, kind):
    n = 100_000 if kind != "duplicates" else 50_000
    v, g, t = oracle_tree(kind, n)
    q = v[:: 4] if kind != "duplicates" else v
    o_ref, i_ref, d_ref = t.query_knn(q, k)
    bvh = ax.BoundingVolumeHierarchy(dev(v), node_format=node_format)
    idx, off, dist = bvh.query(ax.make_nearest(dev(q), k), return_distances=True)
    np.testing.assert_array_equal(host(off), o_ref)
    np.testing.assert_array_equal(host(idx), i_ref)
    np.testing.assert_array_equal(host(dist).view(np.uint32), d_ref.view(np.uint32))


@pytest.mark.parametrize("node_format", ["fp32", "q16"])
def test_knn_grid_ties(node_format):
    gr = np.arange(20, dtype=np.float32)
    v = np.stack(np.meshgrid(gr, gr, gr, indexing="ij"), -1).reshape(-1, 3).copy()
    q = np.floor(synthetic.uniform_points(3000, 43) * 20).astype(np.float32) + np.float32(0.5)
    q[::2] = np.floor(q[::2])
    t = oracle.Tree(v)
    for k in (7, 27, 64, 100):
        o_ref, i_ref, d_ref = t.query_knn(q, k)
        bvh = ax.BoundingVolumeHierarchy(dev(v), node_format=node_format)
        idx, off, dist = bvh.query(ax.make_nearest(dev(q), k), return_distances=True)
        np.testing.assert_array_equal(host(idx), i_ref)
        np.testing.assert_array_equal(host(dist), d_ref)


def test_knn_large_k_and_k_above_n():
    v, g, t = oracle_tree("uniform", 1000)
    q = synthetic.uniform_points(300, 99)
    bvh = ax.BoundingVolumeHierarchy(dev(v))
    for k in (65, 200, 1000, 1500):
        o_ref, i_ref, d_ref = t.query_knn(q, k)
        idx, off, dist = bvh.query(ax.make_nearest(dev(q), k), return_distances=True)
        np.testing.assert_array_equal(host(off), o_ref)
        np.testing.assert_array_equal(host(idx), i_ref)
        np.testing.assert_array_equal(host(dist), d_ref)
    idx, off = bvh.query(ax.make_nearest(dev(q), 0))
    assert host(off).tolist() == [0] * 301


def test_knn_box_leaves():
    v, g, t = oracle_tree("boxes", 50_000)
    q = synthetic.uniform_points(10_000, 12)
    o_ref, i_ref, d_ref = t.query_knn(q, 8)
    for fmt in ("fp32", "q16"):
        bvh = ax.BoundingVolumeHierarchy(dev(v), geometry="boxes", node_format=fmt)
        idx, off, dist = bvh.query(ax.make_nearest(dev(q), 8), return_distances=True)
        np.testing.assert_array_equal(host(idx), i_ref)
        np.testing.assert_array_equal(host(dist), d_ref)


# ------------------------------------------------------------ reference cases
def test_degenerate_cases(kat):
    for c in kat["degenerate"]["cases"]:
        tree = np.array(c["tree"], np.float32).reshape(-1, 6)
        for fmt in ("fp32", "q16"):
            bvh = ax.BoundingVolumeHierarchy(dev(tree), geometry="boxes", node_format=fmt)
            if c["kind"] == "nearest":
                preds = np.array(c["preds"], np.float64).reshape(-1, 4)
                off, idx = [0], []
                for p in preds:
                    i, o = bvh.query(ax.make_nearest(dev(p[None, :3].astype(np.float32)), int(p[3])))
                    idx += host(i).tolist()
                    off.append(len(idx))
            else:
                if c["kind"] == "spheres":
                    preds = np.array(c["preds"], np.float32).reshape(-1, 4)
                    i, o = bvh.query(ax.intersects(ax.Spheres(dev(preds))))
                else:
                    preds = np.array(c["preds"], np.float32).reshape(-1, 6)
                    i, o = bvh.query(ax.intersects(ax.Boxes(dev(preds))))
                off, idx = host(o), host(i)
            assert_crs_equal_as_sets(off, idx, c["offsets"], c["indices"])


def test_manufactured_solution():
    boxes, *_ = manufactured_grid()
    for fmt in ("fp32", "q16"):
        bvh = ax.BoundingVolumeHierarchy(dev(boxes), geometry="boxes", node_format=fmt)
        for name, q, off, ind in manufactured_queries():
            i, o = bvh.query(ax.intersects(ax.Boxes(dev(q))))
            assert_crs_equal_as_sets(host(o), host(i), off, ind)


def test_stack_capacity(kat):
    b = diagonal_boxes(kat["stack_capacity"]["n"])
    FM = np.finfo(np.float32).max
    q = np.array([[FM, FM, FM, -FM, -FM, -FM], [0, 0, 0, 4096, 4096, 4096]], np.float32)
    for fmt in ("fp32", "q16"):
        bvh = ax.BoundingVolumeHierarchy(dev(b), geometry="boxes", node_format=fmt)
        i, o = bvh.query(ax.intersects(ax.Boxes(dev(q))))
        assert int(host(o)[-1]) == kat["stack_capacity"]["spatial_total"]
        i, o = bvh.query(ax.make_nearest(torch.zeros((1, 3), device="cuda"), 4096))
        assert int(host(o)[-1]) == kat["stack_capacity"]["nearest_total"]
        assert sorted(host(i).tolist()) == list(range(4096))


def test_nearest_collinear(kat):
    a = kat["nearest_collinear"]
    bvh = ax.BoundingVolumeHierarchy(dev(np.array(a["points"], np.float32)))
    i, o, d = bvh.query(ax.make_nearest(dev(np.array([a["query"]], np.float32)), a["k"]),
                        return_distances=True)
    assert host(i).tolist() == a["indices"]
    np.testing.assert_array_equal(host(d), np.array(a["distances"], np.float32))


def test_buffer_policy():
    v, g, t = oracle_tree("uniform", 100_000)
    r = float(np.float32(0.0336778))
    o_ref, i_ref = t.query_spatial(0, spheres_from_points(v, r))
    bvh = ax.BoundingVolumeHierarchy(dev(v))
    pts = dev(v)
    # fits; then rows overflow (second pass) at sizes that split the
    # compaction's 64-wide steps unevenly
    for b in (64, 37, 8, 3, 1):
        i, o = bvh.query(ax.make_intersects(pts, r), ax.TraversalPolicy().setBufferSize(b))
        np.testing.assert_array_equal(host(o), o_ref)
        np.testing.assert_array_equal(host(i), i_ref)
    i, o = bvh.query(ax.make_intersects(pts, r), ax.TraversalPolicy().setBufferSize(-64))
    np.testing.assert_array_equal(host(i), i_ref)
    with pytest.raises(ax.SearchException):  # tstQueryTreeTraversalPolicy.cpp:125-128
        bvh.query(ax.make_intersects(pts, r), ax.TraversalPolicy().setBufferSize(-1))
    # the library keeps working after the error
    i, o = bvh.query(ax.make_intersects(pts, r))
    np.testing.assert_array_equal(host(i), i_ref)


def test_concurrent_queries_on_two_streams():
    """A radius query and a kNN query of one tree from two host threads on two
    streams at once (each stream has its own workspaces)."""
    import threading

    v, g, t = oracle_tree("uniform", 100_000)
    r = np.float32((48.0 / (4.0 * math.pi * 100_000)) ** (1.0 / 3.0))
    q = v[::3]
    o_ref, i_ref = t.query_spatial(0, spheres_from_points(v, r))
    ko_ref, ki_ref, kd_ref = t.query_knn(q, 16)
    bvh = ax.BoundingVolumeHierarchy(dev(v), node_format="q16")
    pts, qd = dev(v), dev(q)
    torch.cuda.synchronize()
    out = {}

    def radius(s):
        with torch.cuda.stream(s):
            for it in range(3):
                out[("r", it)] = bvh.query(ax.make_intersects(pts, float(r)), stream=s)

    def knn(s):
        with torch.cuda.stream(s):
            for it in range(3):
                out[("k", it)] = bvh.query(ax.make_nearest(qd, 16), return_distances=True,
                                           stream=s)

    th = [threading.Thread(target=radius, args=(torch.cuda.Stream(),)),
          threading.Thread(target=knn, args=(torch.cuda.Stream(),))]
    for x in th:
        x.start()
    for x in th:
        x.join()
    torch.cuda.synchronize()
    for it in range(3):
        idx, off = out[("r", it)]
        np.testing.assert_array_equal(host(off), o_ref)
        np.testing.assert_array_equal(host(idx), i_ref)
        ki, ko, kd = out[("k", it)]
        np.testing.assert_array_equal(host(ki), ki_ref)
        np.testing.assert_array_equal(host(kd), kd_ref)


# ------------------------------------------------------------ energy
def _fcc_csr(cells, rc=2.5):
    x = synthetic.fcc_lattice(cells, seed=6)
    bvh = ax.BoundingVolumeHierarchy(dev(x))
    idx, off = bvh.query(ax.make_intersects(dev(x), rc))
    t = oracle.Tree(x)
    o_ref, i_ref = t.query_spatial(0, spheres_from_points(x, np.float32(rc)))
    np.testing.assert_array_equal(host(off), o_ref)
    np.testing.assert_array_equal(host(idx), i_ref)
    return x, off, idx


def test_pair_energy_forward_backward_vs_fp64():
    rc = 2.5
    x, off, idx = _fcc_csr((12, 12, 12), rc)
    e_ref, g_ref = oracle.pair_energy_f64(x, host(off), host(idx), rc)
    e, g = ax.pair_energy(dev(x), off, idx, rc)
    e = float(host(e))
    g = host(g).astype(np.float64)
    assert abs(e - e_ref) <= 1e-6 * abs(e_ref)
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    assert rel <= 1e-5, rel
    # energy-only and fused runs agree
    e2, _ = ax.pair_energy(dev(x), off, idx, rc, want_grad=False)
    assert float(host(e2)) == e


def test_pair_energy_fd_subregion():
    rc = 2.5
    x, off, idx = _fcc_csr((10, 10, 10), rc)
    _, g_ref = oracle.pair_energy_f64(x, host(off), host(idx), rc)
    atoms = np.arange(0, x.shape[0], 4)[:1000].astype(np.int32)
    fd = host(ax.pair_energy_fd(dev(x), off, idx, rc, atoms, h=1e-6))
    _, g = ax.pair_energy(dev(x), off, idx, rc, want_energy=False)
    g = host(g).astype(np.float64)[atoms]
    scale = np.abs(g_ref[atoms]).max()
    assert np.abs(fd - g_ref[atoms]).max() <= 1e-6 * scale
    assert np.linalg.norm(g - fd) / np.linalg.norm(fd) <= 1e-5


def test_pair_energy_closed_forms():
    """The GPU energy against closed forms, not only against the oracle: a
    dimer (E = phi(r) - phi(rc), dE/dx_0 = -phi'(r) along the axis) and the
    interior atoms of a perfect FCC crystal (E per atom = 1/2 sum over the
    six shells n_s (phi(r_s) - phi(rc)), zero gradient).  Float pair terms,
    double sums: relative tolerance 1e-6 on E."""
    from tests.test_oracle import fcc_integer_block, lj
    rc = 2.5
    r = 1.5
    x = np.array([[0.25, 0.5, 0.75], [0.25 + r, 0.5, 0.75]], np.float32)
    bvh = ax.BoundingVolumeHierarchy(dev(x))
    idx, off = bvh.query(ax.make_intersects(dev(x), rc))
    e, g = ax.pair_energy(dev(x), off, idx, rc)
    phi, dphi = lj(r)
    phic, _ = lj(rc)
    assert float(host(e)) == pytest.approx(phi - phic, rel=1e-6)
    np.testing.assert_allclose(host(g)[0], [-dphi, 0, 0], rtol=1e-5, atol=1e-6)
    x = fcc_integer_block(7)
    rcf = float(np.float32(2.5 * math.sqrt(2.0)))
    bvh = ax.BoundingVolumeHierarchy(dev(x))
    idx, off = bvh.query(ax.make_intersects(dev(x), rcf))
    off_h, idx_h = host(off), host(idx)
    interior = np.nonzero(np.abs(x - 7.0).max(1) <= 2.0)[0]
    cnt = np.zeros(x.shape[0], np.int64)
    cnt[interior] = np.diff(off_h)[interior]
    o2 = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    i2 = np.concatenate([idx_h[off_h[i]:off_h[i + 1]] for i in interior]).astype(np.int32)
    e, g = ax.pair_energy(dev(x), dev(o2), dev(i2), rcf)
    phic, _ = lj(rcf)
    shells = {2: 12, 4: 6, 6: 24, 8: 12, 10: 24, 12: 8}
    per_atom = 0.5 * sum(n * (lj(math.sqrt(s))[0] - phic) for s, n in shells.items())
    assert float(host(e)) == pytest.approx(len(interior) * per_atom, rel=1e-6)
    assert np.abs(host(g)[interior]).max() <= 1e-5


@pytest.mark.parametrize("recompute", [False, True])
def test_pair_energy_autograd(recompute):
    rc = 2.5
    x = dev(synthetic.fcc_lattice((8, 8, 8), seed=6)).requires_grad_(True)
    e = ax.PairEnergy.apply(x, rc, None, None, recompute)
    e.backward()
    off, idx = ax.energy.neighbour_csr(x.detach(), rc)
    _, g = ax.pair_energy(x.detach(), off, idx, rc, want_energy=False)
    np.testing.assert_array_equal(host(x.grad), host(g))
