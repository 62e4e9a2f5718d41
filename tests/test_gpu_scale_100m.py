"""One step of the 100M-point moving-points loop under a check: 100M points
uniform in [0,1)^3 (seed 7), the Q16 and FP32 builds, kNN k = 32 over all
100M queries, and the full neighbour list + pair energy forward/backward.

The oracle cannot hold the whole step in a test's time, so it is checked in
two ways that together pin it:

* the sort over all 100M points by what fixes a stable sort uniquely: the
  keys are non-decreasing, each key is the oracle's Morton code of the point
  the permutation puts there, the permutation is a permutation, and equal
  keys keep increasing indices;
* rows on an inner cube of about 100k queries against the oracle run on the
  points of a cube 0.02 wider on every side: a query's k nearest points
  (and its neighbours within the cutoff) lie within its k-th distance (the
  cutoff), so when that is below the query's distance to the outer cube's
  faces -- asserted -- the rows over the sub-cloud are the rows over all
  100M points."""
import numpy as np
import pytest
import torch

import arborx_amd as ax
import oracle
from arborx_amd import synthetic

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

N = 100_000_000
K = 32
INNER = (0.45, 0.55)
MARGIN = 0.02

_state = {}


@pytest.fixture(scope="module", autouse=True)
def _release():
    yield
    _state.clear()
    ax.trim_memory()


def points():
    if "pts" not in _state:
        p = synthetic.uniform_points(N, 7)
        _state["pts"] = p
        _state["dev"] = torch.from_numpy(p).cuda()
    return _state["pts"], _state["dev"]


def host(t):
    return t.detach().cpu().numpy()


def in_cube(p, lo, hi):
    return np.nonzero(np.all((p >= lo) & (p < hi), axis=1))[0]


def test_build_at_100m_fixes_the_stable_sort():
    p, d = points()
    bvh = ax.BoundingVolumeHierarchy(d, node_format="q16", keep_topology=True)
    scene = oracle.scene_box(p)
    np.testing.assert_array_equal(np.array(bvh.bounds(), np.float32), scene)
    e = bvh.export_topology()
    keys = host(e["keys"]).view(np.uint64)
    perm = host(e["perm"]).view(np.uint32)
    del e, bvh
    assert np.all(keys[1:] >= keys[:-1])
    seen = np.zeros(N, bool)
    seen[perm] = True
    assert seen.all()
    del seen
    np.testing.assert_array_equal(keys, oracle.morton64_values(p[perm], scene))
    same = keys[1:] == keys[:-1]
    assert np.all(perm[1:][same] > perm[:-1][same])


@pytest.mark.parametrize("fmt", ["q16", "fp32"])
def test_knn_k32_at_100m_on_an_inner_cube(fmt):
    p, d = points()
    bvh = ax.BoundingVolumeHierarchy(d, node_format=fmt)
    ki, _, kd = bvh.query(ax.make_nearest(d, K), return_distances=True)
    del bvh
    sample = in_cube(p, *INNER)
    assert 90_000 < sample.size < 110_000
    s_dev = torch.from_numpy(sample.astype(np.int64)).cuda()
    gi = host(ki.view(N, K)[s_dev])
    gd = host(kd.view(N, K)[s_dev])
    del ki, kd
    sub = in_cube(p, INNER[0] - MARGIN, INNER[1] + MARGIN)
    t = oracle.Tree(p[sub])
    _, oi, od = t.query_knn(p[sample], K)
    oi = sub[oi].reshape(-1, K)
    od = od.reshape(-1, K)
    assert od[:, -1].max() < MARGIN  # every k-th neighbour lies inside the sub-cloud
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))


def test_pair_energy_at_100m_on_an_inner_cube():
    """The loop's energy: coordinates scaled by 2.5 / r32 (r32 = radius of
    about 32 neighbours), cutoff 2.5, the full neighbour list recomputed.
    Rows, gradient and energy of the inner cube's atoms against the oracle
    (FP64): rows identical as sets, gradient relative error <= 1e-5, E of
    those rows relative error <= 1e-6."""
    p, d = points()
    rc = 2.5
    scale = np.float32(rc / synthetic.radius_for_neighbours(N, 32))
    xs = p * scale
    xd = d * float(scale)
    assert torch.equal(xd.cpu(), torch.from_numpy(xs))
    off, idx = ax.find_full_neighbor_list(xd, rc)
    _, g = ax.pair_energy(xd, off, idx, rc, want_energy=False)
    sample = in_cube(p, *INNER)
    s_dev = torch.from_numpy(sample.astype(np.int64)).cuda()
    g_s = host(g[s_dev]).astype(np.float64)
    o_h = host(off)
    lens = o_h[sample + 1] - o_h[sample]
    starts = o_h[sample]
    del g
    # GPU rows of the sample atoms, as a CRS over global indices
    pos = np.repeat(starts - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens) + \
        np.arange(lens.sum())
    rows_g = host(idx[torch.from_numpy(pos).cuda()])
    del off, idx
    # the oracle's rows over the sub-cloud
    sub = in_cube(p, INNER[0] - MARGIN, INNER[1] + MARGIN)
    assert rc / scale < MARGIN
    o_ref, i_ref = oracle.neighbor_list(xs[sub], rc, True)
    where = np.searchsorted(sub, sample)
    r_lens = o_ref[where + 1] - o_ref[where]
    np.testing.assert_array_equal(r_lens, lens)
    r_pos = np.repeat(o_ref[where] - np.concatenate([[0], np.cumsum(r_lens)[:-1]]), r_lens) + \
        np.arange(r_lens.sum())
    rows_o = sub[i_ref[r_pos]]
    cut = np.concatenate([[0], np.cumsum(lens)])
    for a, b in zip(cut[:-1], cut[1:]):
        if not np.array_equal(np.sort(rows_g[a:b]), np.sort(rows_o[a:b])):
            raise AssertionError("row mismatch")
    # FP64 gradient and energy of the sample rows; rows are complete, so
    # each atom's gradient depends on its own row only
    loc = np.searchsorted(sub, rows_o).astype(np.int32)
    o_s = np.zeros(sub.size + 1, np.int64)
    cnt = np.zeros(sub.size, np.int64)
    cnt[where] = r_lens
    o_s[1:] = np.cumsum(cnt)
    order = np.argsort(where, kind="stable")
    i_s = np.concatenate([loc[cut[j]:cut[j + 1]] for j in order]).astype(np.int32)
    e_ref, g_ref = oracle.pair_energy_f64(xs[sub], o_s, i_s, rc)
    g_ref = g_ref[where]
    rel = np.linalg.norm(g_s - g_ref) / np.linalg.norm(g_ref)
    assert rel <= 1e-5, rel
    # the GPU energy of the same rows
    o_gs = np.zeros(sub.size + 1, np.int64)
    o_gs[1:] = np.cumsum(cnt)
    e_gpu, _ = ax.pair_energy(torch.from_numpy(xs[sub]).cuda(), torch.from_numpy(o_gs).cuda(),
                              torch.from_numpy(i_s).cuda(), rc, want_grad=False)
    assert float(host(e_gpu)) == pytest.approx(e_ref, rel=1e-6)
