"""CPU oracle for the ArborX BVH hot path -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper over ``oracle/_build/liboracle.so`` (built from ``oracle.cpp``
by ``oracle/Makefile``), a plain C++/OpenMP restatement of ArborX's Serial
path: scene box and Morton codes (TreeConstruction.hpp:26-39,
SpaceFillingCurves.hpp:45-84, MortonCode.hpp:64-72,188-197,291-347), stable
key sort (SortUtils.hpp:26-43), the Apetrei/Karras climb with ropes
(TreeConstruction.hpp:71-322), rope traversal (TreeTraversal.hpp:97-119),
two-pass CRS (CrsGraphWrapperImpl.hpp:145-374) and stack kNN
(TreeTraversal.hpp:180-335, ties broken by (distance, index)).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this package, and only as the checker or the CPU baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

P = ctypes.c_void_p
I64 = ctypes.c_int64
I32 = ctypes.c_int
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64


def build():
    """Compile liboracle.so with the committed Makefile (g++, no FMA)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _ptr(a):
    return a.ctypes.data_as(P) if a is not None else None


# Predicate kinds of the ordered traversal (those of the C ABI) and their
# floats per predicate.
ORDERED_RAYS = 4
ORDERED_SPHERES = 5
ORDERED_BOXES = 6
ORDERED_POINTS = 7
ORDERED_STRIDE = {ORDERED_RAYS: 6, ORDERED_SPHERES: 4, ORDERED_BOXES: 6, ORDERED_POINTS: 3}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_expand2_u32": (U32, [U32]),
            "orc_expand2_u64": (U64, [U64]),
            "orc_morton32": (U32, [P]),
            "orc_morton64": (U64, [P]),
            "orc_scene_box": (None, [P, I64, I32, P]),
            "orc_morton64_values": (None, [P, I64, I32, P, P]),
            "orc_morton32_values": (None, [P, I64, I32, I32, P, P]),
            "orc_sort_u64": (None, [P, I64, P]),
            "orc_stable_sort_u64": (None, [P, I64, P]),
            "orc_build": (P, [P, I64, I32]),
            "orc_build_from_codes": (P, [P, P, I64]),
            "orc_free": (None, [P]),
            "orc_size": (I64, [P]),
            "orc_bounds": (None, [P, P]),
            "orc_export": (None, [P, P, P, P, P, P, P]),
            "orc_q16_encode_box": (None, [P, P, P, P]),
            "orc_linearize": (I64, [P, I32, I32, P]),
            "orc_query_spatial": (I64, [P, I32, P, I64, P, P, P, P]),
            "orc_bruteforce_spatial": (I64, [P, I64, I32, I32, P, I64, P, P]),
            "orc_query_knn": (I64, [P, P, I64, I32, P, P, P, P, P]),
            "orc_bruteforce_knn": (None, [P, I64, I32, P, I64, I32, P, P]),
            "orc_make_rays": (I64, [P, P, I64, P]),
            "orc_bruteforce_rays_ordered": (I64, [P, I64, P, I64, I64, P, P]),
            "orc_query_ordered": (I64, [P, P, I64, P, P, P]),
            "orc_query_ordered_kind": (I64, [P, I32, P, I64, P, P, P]),
            "orc_bruteforce_ordered": (I64, [P, I64, I32, I32, P, I64, I64, P, P]),
            "orc_ordered_distance": (ctypes.c_float, [I32, P, I32, P]),
            "orc_pair_energy_f64": (ctypes.c_double, [P, I64, P, P, ctypes.c_double, P]),
            "orc_neighbor_list": (I64, [P, I64, ctypes.c_float, I32, P, P]),
            "orc_expand_half_to_full": (I64, [P, P, I64, P, P]),
            "orc_num_threads": (I32, []),
            "orc_set_num_threads": (None, [I32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _f32(a, cols):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.reshape(-1, cols)


def expand2_u32(x):
    return int(lib().orc_expand2_u32(x))


def expand2_u64(x):
    return int(lib().orc_expand2_u64(x))


def morton32(p):
    p = np.ascontiguousarray(p, dtype=np.float32)
    return int(lib().orc_morton32(_ptr(p)))


def morton64(p):
    p = np.ascontiguousarray(p, dtype=np.float32)
    return int(lib().orc_morton64(_ptr(p)))


def scene_box(values, geometry=0):
    v = _f32(values, 3 if geometry == 0 else 6)
    out = np.empty(6, np.float32)
    lib().orc_scene_box(_ptr(v), v.shape[0], geometry, _ptr(out))
    return out


def morton64_values(values, scene, geometry=0):
    v = _f32(values, 3 if geometry == 0 else 6)
    s = np.ascontiguousarray(scene, dtype=np.float32)
    out = np.empty(v.shape[0], np.uint64)
    lib().orc_morton64_values(_ptr(v), v.shape[0], geometry, _ptr(s), _ptr(out))
    return out


def morton32_values(values, scene, stride, geometry):
    v = np.ascontiguousarray(values, dtype=np.float32).reshape(-1, stride)
    s = np.ascontiguousarray(scene, dtype=np.float32)
    out = np.empty(v.shape[0], np.uint32)
    lib().orc_morton32_values(_ptr(v), v.shape[0], stride, geometry, _ptr(s), _ptr(out))
    return out


def sort_u64(keys, stable_sort=False):
    """Returns (sorted keys, permutation of original indices)."""
    k = np.array(keys, dtype=np.uint64, copy=True)
    perm = np.empty(k.shape[0], np.uint32)
    f = lib().orc_stable_sort_u64 if stable_sort else lib().orc_sort_u64
    f(_ptr(k), k.shape[0], _ptr(perm))
    return k, perm


class Tree:
    """Reference-form BVH (Node.hpp layout) built on the CPU."""

    def __init__(self, values, geometry=0, sorted_codes=None):
        self.geometry = geometry
        self._values = _f32(values, 3 if geometry == 0 else 6)
        self.n = self._values.shape[0]
        if sorted_codes is None:
            self._h = lib().orc_build(_ptr(self._values), self.n, geometry)
        else:
            self._codes = np.ascontiguousarray(sorted_codes, dtype=np.uint64)
            assert geometry == 0 and self._codes.shape[0] == self.n
            self._h = lib().orc_build_from_codes(_ptr(self._codes), _ptr(self._values), self.n)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_free(self._h)
            self._h = None

    def bounds(self):
        out = np.empty(6, np.float32)
        lib().orc_bounds(self._h, _ptr(out))
        return out

    def export(self):
        n = self.n
        m = max(n - 1, 0)
        keys = np.zeros(n, np.uint64)
        perm = np.zeros(n, np.uint32)
        leaf_rope = np.zeros(n, np.int32)
        left_child = np.zeros(m, np.int32)
        rope = np.zeros(m, np.int32)
        boxes = np.zeros((m, 6), np.float32)
        lib().orc_export(self._h, _ptr(keys), _ptr(perm), _ptr(leaf_rope),
                         _ptr(left_child), _ptr(rope), _ptr(boxes))
        return dict(keys=keys, perm=perm, leaf_rope=leaf_rope,
                    left_child=left_child, rope=rope, boxes=boxes)

    def linearize(self, node_format=0, flat_max=1):
        """Pre-order slot array (uint32, shape (slots, 4)).  Subtrees of at
        most `flat_max` leaves are stored as their leaves only (1: every
        internal node is stored)."""
        n = self.n
        s_leaf = 1 if self.geometry == 0 else 2
        s_int = 2 if node_format == 0 else 1
        cap = max(n * s_leaf + max(n - 1, 0) * s_int, 1)
        words = np.zeros((cap, 4), np.uint32)
        slots = lib().orc_linearize(self._h, node_format, int(flat_max), _ptr(words))
        return words[:slots]

    def query_ordered(self, rays, kind=ORDERED_RAYS):
        """The reference's ordered (best-first) traversal -> (offsets,
        indices, distances), rows in traversal order.  kind: ORDERED_RAYS
        (nq, 6) {origin, direction}, ORDERED_SPHERES (nq, 4), ORDERED_BOXES
        (nq, 6) or ORDERED_POINTS (nq, 3)."""
        r = _f32(rays, ORDERED_STRIDE[kind])
        nq = r.shape[0]
        offsets = np.zeros(nq + 1, np.int64)
        nnz = lib().orc_query_ordered_kind(self._h, kind, _ptr(r), nq, _ptr(offsets), None,
                                           None)
        if nnz < 0:
            raise RuntimeError("ordered traversal: heap of 64 entries exceeded")
        idx = np.zeros(max(nnz, 1), np.int32)
        dist = np.zeros(max(nnz, 1), np.float32)
        lib().orc_query_ordered_kind(self._h, kind, _ptr(r), nq, _ptr(offsets), _ptr(idx),
                                     _ptr(dist))
        return offsets, idx[:nnz], dist[:nnz]

    def query_spatial(self, kind, preds, with_visits=False):
        """kind 0: spheres (nq,4) {x,y,z,r}; kind 1: boxes (nq,6); kind 3:
        rays (nq,6) {origin, unit direction}."""
        p = _f32(preds, 4 if kind == 0 else 6)
        nq = p.shape[0]
        offsets = np.zeros(nq + 1, np.int64)
        vi = np.zeros(nq, np.int64)
        vl = np.zeros(nq, np.int64)
        nnz = lib().orc_query_spatial(self._h, kind, _ptr(p), nq, _ptr(offsets),
                                      None, None, None)
        indices = np.zeros(max(nnz, 1), np.int32)
        lib().orc_query_spatial(self._h, kind, _ptr(p), nq, _ptr(offsets),
                                _ptr(indices), _ptr(vi), _ptr(vl))
        if with_visits:
            return offsets, indices[:nnz], vi, vl
        return offsets, indices[:nnz]

    def query_knn(self, points, k, with_visits=False):
        p = _f32(points, 3)
        nq = p.shape[0]
        per = 0 if (k < 1 or self.n == 0) else min(k, self.n)
        offsets = np.zeros(nq + 1, np.int64)
        idx = np.zeros(max(nq * per, 1), np.int32)
        dist = np.zeros(max(nq * per, 1), np.float32)
        vi = np.zeros(nq, np.int64)
        vl = np.zeros(nq, np.int64)
        lib().orc_query_knn(self._h, _ptr(p), nq, k, _ptr(offsets), _ptr(idx),
                            _ptr(dist), _ptr(vi), _ptr(vl))
        r = (offsets, idx[: nq * per], dist[: nq * per])
        return r + (vi, vl) if with_visits else r


def bruteforce_spatial(values, geometry, kind, preds):
    v = _f32(values, 3 if geometry == 0 else 6)
    p = _f32(preds, 4 if kind == 0 else 6)
    nq = p.shape[0]
    offsets = np.zeros(nq + 1, np.int64)
    nnz = lib().orc_bruteforce_spatial(_ptr(v), v.shape[0], geometry, kind, _ptr(p), nq,
                                       _ptr(offsets), None)
    idx = np.zeros(max(nnz, 1), np.int32)
    lib().orc_bruteforce_spatial(_ptr(v), v.shape[0], geometry, kind, _ptr(p), nq,
                                 _ptr(offsets), _ptr(idx))
    return offsets, idx[:nnz]


def bruteforce_knn(values, geometry, points, k):
    v = _f32(values, 3 if geometry == 0 else 6)
    p = _f32(points, 3)
    n, nq = v.shape[0], p.shape[0]
    per = 0 if (k < 1 or n == 0) else min(k, n)
    idx = np.zeros(max(nq * per, 1), np.int32)
    dist = np.zeros(max(nq * per, 1), np.float32)
    lib().orc_bruteforce_knn(_ptr(v), n, geometry, _ptr(p), nq, k, _ptr(idx), _ptr(dist))
    return idx[: nq * per], dist[: nq * per]


def make_rays(origins, directions):
    """Rays {origin, unit direction} as the reference's Ray constructor makes
    them (double-precision normalisation); zero components stored as +0."""
    o = _f32(origins, 3)
    d = _f32(directions, 3)
    out = np.zeros((o.shape[0], 6), np.float32)
    bad = lib().orc_make_rays(_ptr(o), _ptr(d), o.shape[0], _ptr(out))
    if bad:
        raise ValueError("make_rays: zero direction")
    return out


def bruteforce_ordered(values, geometry, kind, preds, limit=-1):
    """Rows of the values at a finite distance from each ordered predicate
    (kind ORDERED_RAYS / _SPHERES / _BOXES / _POINTS), in (distance, index)
    order, at most `limit` per row (-1: all) -> (offsets, indices)."""
    v = _f32(values, 3 if geometry == 0 else 6)
    q = _f32(preds, ORDERED_STRIDE[kind])
    nq = q.shape[0]
    offsets = np.zeros(nq + 1, np.int64)
    nnz = lib().orc_bruteforce_ordered(_ptr(v), v.shape[0], geometry, kind, _ptr(q), nq, limit,
                                       _ptr(offsets), None)
    idx = np.zeros(max(nnz, 1), np.int32)
    lib().orc_bruteforce_ordered(_ptr(v), v.shape[0], geometry, kind, _ptr(q), nq, limit,
                                 _ptr(offsets), _ptr(idx))
    return offsets, idx[:nnz]


def ordered_distance(kind, pred, geometry, value):
    """distance(predicate geometry, value) of the ordered traversal."""
    q = np.ascontiguousarray(pred, dtype=np.float32)
    v = np.ascontiguousarray(value, dtype=np.float32)
    return float(lib().orc_ordered_distance(kind, _ptr(q), geometry, _ptr(v)))


def bruteforce_rays_ordered(boxes, rays, limit=-1):
    """Rows of the boxes each ray hits, in (distance, index) order, at most
    `limit` per row (-1: all) -> (offsets, indices)."""
    b = _f32(boxes, 6)
    r = _f32(rays, 6)
    nq = r.shape[0]
    offsets = np.zeros(nq + 1, np.int64)
    nnz = lib().orc_bruteforce_rays_ordered(_ptr(b), b.shape[0], _ptr(r), nq, limit,
                                            _ptr(offsets), None)
    idx = np.zeros(max(nnz, 1), np.int32)
    lib().orc_bruteforce_rays_ordered(_ptr(b), b.shape[0], _ptr(r), nq, limit, _ptr(offsets),
                                      _ptr(idx))
    return offsets, idx[:nnz]


def q16_encode_box(frame6, box6):
    f = np.ascontiguousarray(frame6, dtype=np.float32)
    b = np.ascontiguousarray(box6, dtype=np.float32)
    q = np.zeros(6, np.uint16)
    dec = np.zeros(6, np.float32)
    lib().orc_q16_encode_box(_ptr(f), _ptr(b), _ptr(q), _ptr(dec))
    return q, dec


def pair_energy_f64(xyz, offsets, indices, rc, with_grad=True):
    x = _f32(xyz, 3)
    o = np.ascontiguousarray(offsets, dtype=np.int64)
    i = np.ascontiguousarray(indices, dtype=np.int32)
    g = np.zeros((x.shape[0], 3), np.float64) if with_grad else None
    e = lib().orc_pair_energy_f64(_ptr(x), x.shape[0], _ptr(o), _ptr(i), float(rc), _ptr(g))
    return (e, g) if with_grad else e


def neighbor_list(points, r, full):
    """findFullNeighborList / findHalfNeighborList (NeighborList.hpp:49-188):
    (offsets int64 (n+1,), indices int32).  Row order is the reference's
    serial order; compare rows as multisets."""
    x = _f32(points, 3)
    n = x.shape[0]
    off = np.zeros(n + 1, np.int64)
    nnz = lib().orc_neighbor_list(_ptr(x), n, float(np.float32(r)), 1 if full else 0,
                                  _ptr(off), None)
    idx = np.zeros(nnz, np.int32)
    lib().orc_neighbor_list(_ptr(x), n, float(np.float32(r)), 1 if full else 0, _ptr(off),
                            _ptr(idx))
    return off, idx


def expand_half_to_full(offsets, indices):
    """expandHalfToFull (ExpandHalfToFull.hpp:23-68) -> (offsets, indices)."""
    o = np.ascontiguousarray(offsets, np.int64)
    i = np.ascontiguousarray(indices, np.int32)
    n = o.shape[0] - 1
    out_o = np.zeros(n + 1, np.int64)
    m = lib().orc_expand_half_to_full(_ptr(o), _ptr(i), n, _ptr(out_o), None)
    out_i = np.zeros(m, np.int32)
    lib().orc_expand_half_to_full(_ptr(o), _ptr(i), n, _ptr(out_o), _ptr(out_i))
    return out_o, out_i


def num_threads():
    return int(lib().orc_num_threads())


def set_num_threads(n):
    lib().orc_set_num_threads(int(n))
