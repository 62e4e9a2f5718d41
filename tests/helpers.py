"""Shared test helpers: CRS comparison and the reference's manufactured
solution (test/tstQueryTreeManufacturedSolution.cpp:29-270) restated."""
import numpy as np


def rows(offsets, indices):
    offsets = np.asarray(offsets, dtype=np.int64)
    indices = np.asarray(indices)
    return [indices[offsets[i]:offsets[i + 1]] for i in range(len(offsets) - 1)]


def assert_crs_equal_as_sets(off_a, ind_a, off_b, ind_b):
    """Rows compared as multisets (test/boost_ext/CompressedStorageComparison.hpp:21-61)."""
    off_a = np.asarray(off_a, dtype=np.int64)
    off_b = np.asarray(off_b, dtype=np.int64)
    assert off_a.shape == off_b.shape, (off_a.shape, off_b.shape)
    np.testing.assert_array_equal(off_a, off_b)
    ra, rb = rows(off_a, ind_a), rows(off_b, ind_b)
    for i, (x, y) in enumerate(zip(ra, rb)):
        if not np.array_equal(np.sort(x), np.sort(y)):
            raise AssertionError(f"row {i}: {np.sort(x)[:20]} != {np.sort(y)[:20]}")


def manufactured_grid():
    """11^3 structured grid of degenerate boxes, spacing 10 (float)."""
    nx = ny = nz = 11
    h = np.float32(100.0) / np.float32(10)
    ind = lambda i, j, k: i + j * nx + k * nx * ny  # noqa: E731
    n = nx * ny * nz
    boxes = np.zeros((n, 6), np.float32)
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                p = [np.float32(i) * h, np.float32(j) * h, np.float32(k) * h]
                boxes[ind(i, j, k)] = p + p
    return boxes, (nx, ny, nz), h, ind


def manufactured_queries():
    """Yields (name, query boxes, offsets, indices) for cases (i)-(iii)."""
    boxes, (nx, ny, nz), h, ind = manufactured_grid()
    n = boxes.shape[0]
    # (i) the objects themselves
    yield "self", boxes.copy(), np.arange(n + 1), np.arange(n)
    # (ii) boxes reaching the first neighbours (27-stencil)
    q = np.zeros((n, 6), np.float32)
    ref = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                q[ind(i, j, k)] = [np.float32(i - 1) * h, np.float32(j - 1) * h,
                                   np.float32(k - 1) * h, np.float32(i + 1) * h,
                                   np.float32(j + 1) * h, np.float32(k + 1) * h]
    for idx in range(n):
        i, j, k = idx % nx, (idx // nx) % ny, idx // (nx * ny)
        s = []
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for dk in (-1, 0, 1):
                    a, b, c = i + di, j + dj, k + dk
                    if 0 <= a < nx and 0 <= b < ny and 0 <= c < nz:
                        s.append(ind(a, b, c))
        ref.append(sorted(s))
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in ref])
    yield "first_neighbours", q, off, np.concatenate([np.array(r) for r in ref])
    # (iii) random boxes around lattice nodes (shift within 0.45 h)
    rng = np.random.default_rng(12345)
    ii = rng.integers(0, nx, n)
    jj = rng.integers(0, ny, n)
    kk = rng.integers(0, nz, n)
    sh = rng.uniform(-0.45, 0.45, (n, 3)).astype(np.float32)
    c = np.stack([(ii + sh[:, 0]) * h, (jj + sh[:, 1]) * h, (kk + sh[:, 2]) * h], 1).astype(np.float32)
    q = np.concatenate([c - h / 2, c + h / 2], 1).astype(np.float32)
    yield "random_points", q, np.arange(n + 1), np.array([ind(a, b, c_) for a, b, c_ in zip(ii, jj, kk)])


def diagonal_boxes(n=4096):
    """tstQueryTreeDegenerate.cpp:375-409: boxes [i, i+1]^3."""
    a = np.arange(n, dtype=np.float32)
    return np.stack([a, a, a, a + 1, a + 1, a + 1], 1).astype(np.float32)


def spheres_from_points(points, r):
    p = np.asarray(points, np.float32)
    return np.concatenate([p, np.full((p.shape[0], 1), np.float32(r), np.float32)], 1)


def ray_inputs(n, seed, boxes=None):
    """Origins and (unnormalised) directions of n test rays: origins in
    [-0.2, 1.2)^3, Gaussian directions, and the awkward cases mixed in:
    directions along an axis (two zero components, some of them -0), origins
    on a face of one of `boxes`, and origins inside a box."""
    g = np.random.default_rng(seed)
    o = g.uniform(-0.2, 1.2, (n, 3)).astype(np.float32)
    d = g.normal(size=(n, 3)).astype(np.float32)
    ax = g.integers(0, 3, n)
    sel = np.arange(n) % 5 == 1
    for i in np.nonzero(sel)[0]:
        keep = d[i, ax[i]]
        d[i] = np.float32(0.0)
        d[i, ax[i]] = keep
        if i % 2:
            d[i, (ax[i] + 1) % 3] = np.float32(-0.0)
    if boxes is not None and len(boxes):
        b = np.asarray(boxes, np.float32).reshape(-1, 6)
        on_face = np.nonzero(np.arange(n) % 7 == 3)[0]
        pick = g.integers(0, len(b), len(on_face))
        for i, j in zip(on_face, pick):
            o[i] = (b[j, :3] + b[j, 3:]) / np.float32(2)
            o[i, ax[i]] = b[j, ax[i]]  # on the lower face along the axis
        inside = np.nonzero(np.arange(n) % 7 == 5)[0]
        pick = g.integers(0, len(b), len(inside))
        o[inside] = (b[pick, :3] + b[pick, 3:]) / np.float32(2)
    return o, d


def canonical_ordered_rows(offsets, indices, distances):
    """Rows of an ordered query with each run of equal distances sorted by
    index; checks that every row is non-decreasing in distance."""
    out = np.array(indices, copy=True)
    offsets = np.asarray(offsets, np.int64)
    for i in range(len(offsets) - 1):
        a, b = offsets[i], offsets[i + 1]
        d = np.asarray(distances[a:b])
        assert np.all(d[1:] >= d[:-1]), f"row {i} not ordered by distance"
        order = np.lexsort((out[a:b], d))
        out[a:b] = out[a:b][order]
    return out
