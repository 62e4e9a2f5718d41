"""arborx_amd -- MI355X (gfx950) BVH: Python mirror of ArborX's BVH interface.

The compute path is the hand-written HIP library ``lib/libarborx_mi355x.so``
reached through its C ABI (``include/arborx_mi355x.h``) with ctypes.  PyTorch
is used only for device memory and streams: tensors are passed by device
pointer, and result arrays the library sizes itself are allocated from
torch's caching allocator through the library's output-allocator hook.

Names and argument meaning follow the reference
(src/spatial/ArborX_LinearBVH.hpp, detail/ArborX_Predicates.hpp,
detail/ArborX_PredicateHelpers.hpp, detail/ArborX_TraversalPolicy.hpp,
misc/ArborX_Exception.hpp):

    bvh = BoundingVolumeHierarchy(points)                 # or boxes
    indices, offsets = bvh.query(make_intersects(points, r))
    indices, offsets = bvh.query(intersects(Boxes(b)), TraversalPolicy())
    indices, offsets, dist = bvh.query(make_nearest(points, k), return_distances=True)

There is no CPU fallback: without the built library (or without a GPU) every
call raises.
"""
import ctypes
import os

import torch  # must load before the library: one HIP runtime per process

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libarborx_mi355x.so")

AXB_OK = 0
AXB_ERR_INVALID_ARGUMENT = -1
AXB_ERR_HIP = -2
AXB_ERR_OUT_OF_MEMORY = -3
AXB_ERR_BUFFER_OVERFLOW = -4
AXB_ERR_INTERNAL = -5

GEOMETRY_POINTS = 0
GEOMETRY_BOXES = 1
NODES_FP32 = 0
NODES_Q16 = 1
BUILD_KEEP_TOPOLOGY = 1
# internal nodes over at most this many leaves are not stored (AXB_FLAT_SUBTREE_MAX)
FLAT_SUBTREE_MAX = 4
PREDICATE_SPHERES = 0
PREDICATE_BOXES = 1
PREDICATE_RAYS = 3
PREDICATE_ORDERED_RAYS = 4
PREDICATE_ORDERED_SPHERES = 5
PREDICATE_ORDERED_BOXES = 6
PREDICATE_ORDERED_POINTS = 7

# Symbols of include/arborx_mi355x.h (checked by the CPU test suite).
EXPORTED_SYMBOLS = (
    "axb_last_error", "axb_version", "axb_set_output_allocator", "axb_free", "axb_trim_memory",
    "axb_bvh_create", "axb_bvh_destroy", "axb_bvh_size", "axb_bvh_bounds",
    "axb_bvh_query_intersects", "axb_bvh_query_intersects_radius",
    "axb_bvh_query_nearest", "axb_bvh_query_nearest_rays", "axb_make_rays",
    "axb_bvh_neighbor_list", "axb_find_neighbor_list",
    "axb_expand_half_to_full", "axb_pair_energy", "axb_pair_energy_noalloc",
    "axb_pair_energy_fd",
    "axb_bruteforce_create", "axb_bruteforce_destroy", "axb_bruteforce_size",
    "axb_bruteforce_bounds", "axb_bruteforce_query_intersects",
    "axb_bruteforce_query_intersects_radius", "axb_bruteforce_query_nearest",
    "axb_bruteforce_query_nearest_rays",
    "axb_bruteforce_view",
    "axb_bvh_view", "axb_bvh_predicate_order", "axb_exclusive_scan_i64", "axb_gather_pairs",
    "axb_bvh_num_slots", "axb_bvh_slots", "axb_bvh_export_topology",
    "axb_sqrt_threshold", "axb_morton64_host", "axb_q16_encode_host", "axb_profile_enable",
    "axb_profile_clear",
    "axb_profile_num_records", "axb_profile_record", "axb_profile_visits",
    "axb_debug_set_limits",
)


class SearchException(RuntimeError):
    """ArborX::SearchException (misc/ArborX_Exception.hpp:19-37)."""


class _Policy(ctypes.Structure):
    _fields_ = [("buffer_size", ctypes.c_int32), ("sort_predicates", ctypes.c_int32)]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_ALLOC_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p)
_FREE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p)

_lib = None
_live_outputs = {}  # device address -> torch tensor owning it


def _torch_alloc(nbytes, _ctx):
    t = torch.empty(int(nbytes), dtype=torch.uint8, device="cuda")
    _live_outputs[t.data_ptr()] = t
    return t.data_ptr()


def _torch_free(ptr, _ctx):
    _live_outputs.pop(ptr, None)


_alloc_cb = _ALLOC_FN(_torch_alloc)
_free_cb = _FREE_FN(_torch_free)


def lib():
    """The loaded C-ABI library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m arborx_amd.build` "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "axb_last_error": (ctypes.c_char_p, []),
            "axb_version": (ctypes.c_char_p, []),
            "axb_set_output_allocator": (_I32, [_ALLOC_FN, _FREE_FN, _P]),
            "axb_free": (_I32, [_P, _P]),
            "axb_trim_memory": (_I32, []),
            "axb_bvh_create": (_I32, [_P, _I64, _I32, _I32, _I32, _P, ctypes.POINTER(_P)]),
            "axb_bvh_destroy": (None, [_P]),
            "axb_bvh_size": (_I64, [_P]),
            "axb_bvh_bounds": (_I32, [_P, _P]),
            "axb_bvh_query_intersects": (_I32, [_P, _I32, _P, _I64, _Policy, _P,
                                                ctypes.POINTER(_P), ctypes.POINTER(_I64), _P]),
            "axb_bvh_query_intersects_radius": (_I32, [_P, _P, _I64, ctypes.c_float, _Policy,
                                                       _P, ctypes.POINTER(_P),
                                                       ctypes.POINTER(_I64), _P]),
            "axb_bvh_query_nearest": (_I32, [_P, _P, _I64, _I32, _Policy, _P,
                                             ctypes.POINTER(_P), ctypes.POINTER(_P),
                                             ctypes.POINTER(_I64), _P]),
            "axb_bvh_query_nearest_rays": (_I32, [_P, _P, _I64, _I32, _Policy, _P,
                                                  ctypes.POINTER(_P), ctypes.POINTER(_I64),
                                                  _P]),
            "axb_make_rays": (_I32, [_P, _P, _I64, _P, _P]),
            "axb_bvh_neighbor_list": (_I32, [_P, ctypes.c_float, _I32, _Policy, _P,
                                             ctypes.POINTER(_P), ctypes.POINTER(_I64), _P]),
            "axb_find_neighbor_list": (_I32, [_P, _I64, ctypes.c_float, _I32, _P,
                                              ctypes.POINTER(_P), ctypes.POINTER(_I64), _P]),
            "axb_expand_half_to_full": (_I32, [_P, _P, _I64, _P, ctypes.POINTER(_P),
                                               ctypes.POINTER(_I64), _P]),
            "axb_bruteforce_create": (_I32, [_P, _I64, _I32, _P, ctypes.POINTER(_P)]),
            "axb_bruteforce_destroy": (None, [_P]),
            "axb_bruteforce_size": (_I64, [_P]),
            "axb_bruteforce_bounds": (_I32, [_P, _P]),
            "axb_bruteforce_query_intersects": (_I32, [_P, _I32, _P, _I64, _Policy, _P,
                                                       ctypes.POINTER(_P), ctypes.POINTER(_I64),
                                                       _P]),
            "axb_bruteforce_query_intersects_radius": (_I32, [_P, _P, _I64, ctypes.c_float,
                                                              _Policy, _P, ctypes.POINTER(_P),
                                                              ctypes.POINTER(_I64), _P]),
            "axb_bruteforce_query_nearest": (_I32, [_P, _P, _I64, _I32, _Policy, _P,
                                                    ctypes.POINTER(_P), ctypes.POINTER(_P),
                                                    ctypes.POINTER(_I64), _P]),
            "axb_bruteforce_query_nearest_rays": (_I32, [_P, _P, _I64, _I32, _Policy, _P,
                                                         ctypes.POINTER(_P),
                                                         ctypes.POINTER(_I64), _P]),
            "axb_pair_energy": (_I32, [_P, _I64, _P, _P, ctypes.c_float, _P, _P, _P]),
            "axb_pair_energy_noalloc": (_I32, [_P, _I64, _P, _P, ctypes.c_float, _P, _P, _P]),
            "axb_pair_energy_fd": (_I32, [_P, _I64, _P, _P, ctypes.c_float, _P, _I64,
                                          ctypes.c_double, _P, _P]),
            "axb_bvh_num_slots": (_I64, [_P]),
            "axb_bvh_slots": (_P, [_P]),
            "axb_bvh_export_topology": (_I32, [_P, _P, _P, _P, _P, _P, _P, _P]),
            "axb_sqrt_threshold": (ctypes.c_float, [ctypes.c_float]),
            "axb_morton64_host": (ctypes.c_uint64, [_P, _P]),
            "axb_q16_encode_host": (None, [_P, _P, _P]),
            "axb_profile_enable": (_I32, [_I32]),
            "axb_profile_clear": (_I32, []),
            "axb_profile_num_records": (_I32, []),
            "axb_profile_record": (_I32, [_I32, ctypes.c_char_p, _I32,
                                          ctypes.POINTER(ctypes.c_double)]),
            "axb_profile_visits": (_I32, [ctypes.POINTER(_I64)]),
            "axb_debug_set_limits": (_I32, [_I64, _I64, _I32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _use_torch_allocator():
    st = lib().axb_set_output_allocator(_alloc_cb, _free_cb, None)
    _check(st)


def _check(status):
    if status == AXB_OK:
        return
    msg = lib().axb_last_error().decode()
    if status == AXB_ERR_BUFFER_OVERFLOW:
        raise SearchException(msg)
    raise RuntimeError(f"arborx_amd error {status}: {msg}")


def _take_output(ptr, count, dtype):
    """Turns a library-allocated output pointer into the owning torch tensor."""
    if not ptr:
        return torch.empty(0, dtype=dtype, device="cuda")
    owner = _live_outputs.pop(ptr, None)
    if owner is None:
        raise RuntimeError("output was not allocated by the torch allocator hook")
    itemsize = torch.empty(0, dtype=dtype).element_size()
    return owner[: count * itemsize].view(dtype)


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _dev_f32(x, cols):
    x = torch.as_tensor(x)
    if x.device.type != "cuda":
        x = x.to("cuda")
    x = x.to(torch.float32).contiguous()
    if x.numel() == 0:
        return x.reshape(0, cols)
    return x.reshape(-1, cols)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else ctypes.c_void_p(0)


# ------------------------------------------------------------ predicates
class Spheres:
    """nq spheres, tensor (nq, 4) {x, y, z, r} (ArborX::Sphere)."""

    def __init__(self, xyzr):
        self.data = _dev_f32(xyzr, 4)


class Boxes:
    """nq boxes, tensor (nq, 6) {lo.xyz, hi.xyz} (ArborX::Box)."""

    def __init__(self, lohi):
        self.data = _dev_f32(lohi, 6)


class Points:
    """nq points, tensor (nq, 3) (ArborX::Point), the geometry of
    ordered_intersects(point)."""

    def __init__(self, xyz):
        self.data = _dev_f32(xyz, 3)


class Rays:
    """nq rays, Experimental::Ray(origin, direction) over a batch
    (src/geometry/ArborX_Ray.hpp:28-66): `data` is a (nq, 6) tensor
    {origin, unit direction}.  The directions are normalised on the device
    as the reference's constructor does (norm and division in double); a
    zero direction raises ValueError.  Rays.from_normalized(t) takes (nq, 6)
    data as is."""

    def __init__(self, origins, directions, stream=None):
        o = _dev_f32(origins, 3)
        d = _dev_f32(directions, 3)
        if o.shape != d.shape:
            raise ValueError("Rays: origins and directions differ in shape")
        self.data = torch.empty((o.shape[0], 6), dtype=torch.float32, device=o.device)
        st = lib().axb_make_rays(_ptr(o), _ptr(d), o.shape[0], _ptr(self.data), _stream(stream))
        if st == AXB_ERR_INVALID_ARGUMENT:
            raise ValueError(lib().axb_last_error().decode())
        _check(st)

    @classmethod
    def from_normalized(cls, data):
        r = cls.__new__(cls)
        r.data = _dev_f32(data, 6)
        return r


class Intersects:
    """ArborX::Intersects<Geometry> over a batch (detail/ArborX_Predicates.hpp:83-101)."""

    def __init__(self, geometry):
        if not isinstance(geometry, (Spheres, Boxes, Rays)):
            raise TypeError("intersects() takes Spheres, Boxes or Rays")
        self.geometry = geometry


class OrderedIntersects:
    """Experimental::OrderedSpatial<Geometry> over a batch
    (detail/ArborX_Predicates.hpp:104-155): for Rays, rows list the boxes
    each ray hits in increasing (distance, index) order; for Spheres, Boxes
    or Points every value is at a finite distance, so each row lists all the
    index's values by (distance(geometry, value), index)."""

    def __init__(self, geometry):
        if not isinstance(geometry, (Rays, Spheres, Boxes, Points)):
            raise TypeError("ordered_intersects() takes Rays, Spheres, Boxes or Points")
        self.geometry = geometry


class _PointsWithRadius:
    def __init__(self, points, r):
        self.points = _dev_f32(points, 3)
        self.r = float(r)


class Nearest:
    """ArborX::Nearest<Point> over a batch with one k (Predicates.hpp:57-81)."""

    def __init__(self, points, k=1):
        self.points = _dev_f32(points, 3)
        self.k = int(k)


class NearestRays:
    """ArborX::Nearest<Ray> over a batch with one k: the first min(k, hits)
    boxes of each ray by (distance, index); rows may be shorter than k."""

    def __init__(self, rays, k=1):
        self.rays = rays
        self.k = int(k)


def intersects(geometry):
    """ArborX::intersects (detail/ArborX_Predicates.hpp:135-140)."""
    return Intersects(geometry)


def nearest(points, k=1):
    """ArborX::nearest (detail/ArborX_Predicates.hpp:129-133); points or Rays."""
    if isinstance(points, Rays):
        return NearestRays(points, k)
    return Nearest(points, k)


def ordered_intersects(geometry):
    """Experimental::ordered_intersects (detail/ArborX_Predicates.hpp:148-155)
    over a batch of Rays, Spheres, Boxes or Points."""
    return OrderedIntersects(geometry)


def make_ordered_intersects(geometry):
    """Experimental::make_ordered_intersects over a batch of Rays, Spheres,
    Boxes or Points."""
    return OrderedIntersects(geometry)


def make_intersects(points, r):
    """Experimental::make_intersects(points, r): spheres of radius r (stored as
    float) around the points (detail/ArborX_PredicateHelpers.hpp:99-105)."""
    if not r > 0:
        raise ValueError("make_intersects: r must be > 0")
    return _PointsWithRadius(points, r)


def make_nearest(points, k):
    """Experimental::make_nearest(points, k) (PredicateHelpers.hpp:114-119);
    points or Rays."""
    return nearest(points, k)


class TraversalPolicy:
    """Experimental::TraversalPolicy (detail/ArborX_TraversalPolicy.hpp:19-47)."""

    def __init__(self, buffer_size=0, sort_predicates=True):
        self._buffer_size = int(buffer_size)
        self._sort_predicates = bool(sort_predicates)

    def setBufferSize(self, buffer_size):
        self._buffer_size = int(buffer_size)
        return self

    def setPredicateSorting(self, sort_predicates):
        self._sort_predicates = bool(sort_predicates)
        return self

    def _c(self):
        return _Policy(self._buffer_size, 1 if self._sort_predicates else 0)


# ------------------------------------------------------------ the tree
class BoundingVolumeHierarchy:
    """ArborX::BoundingVolumeHierarchy over points (N, 3) or boxes (N, 6).

    Leaf values are the indices of `values` (the reference's
    BVH<MemorySpace, int, ...> over Iota / attach_indices form).
    node_format: "fp32" (32-byte internal nodes) or "q16" (16-byte nodes,
    quantised outward; query results are identical).
    """

    def __init__(self, values=None, geometry=None, node_format="fp32",
                 keep_topology=False, stream=None):
        self._h = None
        _use_torch_allocator()
        if values is None:
            values = torch.empty(0, 3)
            geometry = "points"
        v = torch.as_tensor(values)
        if geometry is None:
            geometry = "boxes" if (v.dim() == 2 and v.shape[1] == 6) else "points"
        self.geometry = geometry
        g = GEOMETRY_POINTS if geometry == "points" else GEOMETRY_BOXES
        self.node_format = node_format
        f = NODES_FP32 if node_format == "fp32" else NODES_Q16
        v = _dev_f32(v, 3 if g == GEOMETRY_POINTS else 6)
        h = ctypes.c_void_p()
        flags = BUILD_KEEP_TOPOLOGY if keep_topology else 0
        _check(lib().axb_bvh_create(_ptr(v), v.shape[0], g, f, flags, _stream(stream),
                                    ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.axb_bvh_destroy(self._h)
            self._h = None

    def size(self):
        return int(lib().axb_bvh_size(self._h))

    def empty(self):
        return self.size() == 0

    def bounds(self):
        out = (ctypes.c_float * 6)()
        _check(lib().axb_bvh_bounds(self._h, ctypes.cast(out, _P)))
        return [float(x) for x in out]

    def num_slots(self):
        return int(lib().axb_bvh_num_slots(self._h))

    def slots(self):
        """Copy of the pre-order node array as an int32 tensor (slots, 4)."""
        n = self.num_slots()
        out = torch.empty((n, 4), dtype=torch.int32, device="cuda")
        if n:
            src = lib().axb_bvh_slots(self._h)
            torch.cuda.synchronize()
            _cuda_memcpy_d2d(out, src, n * 16)
        return out

    def export_topology(self, stream=None):
        """Reference-form arrays (tree built with keep_topology=True)."""
        n = self.size()
        m = max(n - 1, 0)
        d = dict(
            keys=torch.zeros(n, dtype=torch.int64, device="cuda"),
            perm=torch.zeros(n, dtype=torch.int32, device="cuda"),
            leaf_rope=torch.zeros(n, dtype=torch.int32, device="cuda"),
            left_child=torch.zeros(m, dtype=torch.int32, device="cuda"),
            rope=torch.zeros(m, dtype=torch.int32, device="cuda"),
            boxes=torch.zeros((m, 6), dtype=torch.float32, device="cuda"),
        )
        _check(lib().axb_bvh_export_topology(
            self._h, _ptr(d["keys"]), _ptr(d["perm"]), _ptr(d["leaf_rope"]),
            _ptr(d["left_child"]), _ptr(d["rope"]), _ptr(d["boxes"]), _stream(stream)))
        torch.cuda.synchronize()
        return d

    def query(self, predicates, policy=None, return_distances=False, stream=None):
        """query(space, predicates, indices, offsets[, policy]) -> (indices, offsets)
        (src/spatial/ArborX_LinearBVH.hpp:90-110).  Nearest predicates with
        return_distances=True also return the float distances."""
        return _crs_query("axb_bvh", self._h, predicates, policy, return_distances, stream)


    def neighbor_list(self, radius, full=True, policy=None, stream=None):
        """Full or half neighbour list of a points tree with itself
        (detail/ArborX_NeighborList.hpp:49-188 semantics, rows by value
        index) -> (offsets, indices)."""
        policy = policy or TraversalPolicy()
        n = self.size()
        offsets = torch.empty(n + 1, dtype=torch.int64, device="cuda")
        ind = ctypes.c_void_p()
        nnz = ctypes.c_int64()
        _check(lib().axb_bvh_neighbor_list(self._h, float(radius), 1 if full else 0,
                                           policy._c(), _ptr(offsets), ctypes.byref(ind),
                                           ctypes.byref(nnz), _stream(stream)))
        return offsets, _take_output(ind.value, nnz.value, torch.int32)


def _crs_query(prefix, handle, predicates, policy, return_distances, stream):
    """The CRS query of an index whose C entry points start with `prefix`."""
    policy = policy or TraversalPolicy()
    s = _stream(stream)
    ind = ctypes.c_void_p()
    nnz = ctypes.c_int64()
    L = lib()
    if isinstance(predicates, Nearest):
        pts = predicates.points
        nq = pts.shape[0]
        offsets = torch.empty(nq + 1, dtype=torch.int64, device="cuda")
        dist = ctypes.c_void_p()
        _check(getattr(L, prefix + "_query_nearest")(
            handle, _ptr(pts), nq, predicates.k, policy._c(), _ptr(offsets), ctypes.byref(ind),
            ctypes.byref(dist) if return_distances else None, ctypes.byref(nnz), s))
        indices = _take_output(ind.value, nnz.value, torch.int32)
        if return_distances:
            return indices, offsets, _take_output(dist.value, nnz.value, torch.float32)
        return indices, offsets
    if isinstance(predicates, _PointsWithRadius):
        pts = predicates.points
        nq = pts.shape[0]
        offsets = torch.empty(nq + 1, dtype=torch.int64, device="cuda")
        _check(getattr(L, prefix + "_query_intersects_radius")(
            handle, _ptr(pts), nq, predicates.r, policy._c(), _ptr(offsets), ctypes.byref(ind),
            ctypes.byref(nnz), s))
        return _take_output(ind.value, nnz.value, torch.int32), offsets
    if isinstance(predicates, NearestRays):
        r = predicates.rays.data
        nq = r.shape[0]
        offsets = torch.empty(nq + 1, dtype=torch.int64, device="cuda")
        _check(getattr(L, prefix + "_query_nearest_rays")(
            handle, _ptr(r), nq, predicates.k, policy._c(), _ptr(offsets), ctypes.byref(ind),
            ctypes.byref(nnz), s))
        return _take_output(ind.value, nnz.value, torch.int32), offsets
    if isinstance(predicates, (Intersects, OrderedIntersects)):
        g = predicates.geometry
        if isinstance(predicates, OrderedIntersects):
            kind = {Rays: PREDICATE_ORDERED_RAYS, Spheres: PREDICATE_ORDERED_SPHERES,
                    Boxes: PREDICATE_ORDERED_BOXES, Points: PREDICATE_ORDERED_POINTS}[type(g)]
        elif isinstance(g, Rays):
            kind = PREDICATE_RAYS
        else:
            kind = PREDICATE_SPHERES if isinstance(g, Spheres) else PREDICATE_BOXES
        nq = g.data.shape[0]
        offsets = torch.empty(nq + 1, dtype=torch.int64, device="cuda")
        _check(getattr(L, prefix + "_query_intersects")(
            handle, kind, _ptr(g.data), nq, policy._c(), _ptr(offsets), ctypes.byref(ind),
            ctypes.byref(nnz), s))
        return _take_output(ind.value, nnz.value, torch.int32), offsets
    raise TypeError("unsupported predicates")


class BruteForce:
    """ArborX::BruteForce over points (N, 3) or boxes (N, 6)
    (src/spatial/ArborX_BruteForce.hpp): the BVH's construction and query
    interface, answered by testing every predicate against every value.
    Spatial rows list values in increasing index order; nearest rows are
    ordered by (distance, index)."""

    def __init__(self, values=None, geometry=None, stream=None):
        self._h = None
        _use_torch_allocator()
        if values is None:
            values = torch.empty(0, 3)
            geometry = "points"
        v = torch.as_tensor(values)
        if geometry is None:
            geometry = "boxes" if (v.dim() == 2 and v.shape[1] == 6) else "points"
        self.geometry = geometry
        g = GEOMETRY_POINTS if geometry == "points" else GEOMETRY_BOXES
        v = _dev_f32(v, 3 if g == GEOMETRY_POINTS else 6)
        h = ctypes.c_void_p()
        _check(lib().axb_bruteforce_create(_ptr(v), v.shape[0], g, _stream(stream),
                                           ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.axb_bruteforce_destroy(self._h)
            self._h = None

    def size(self):
        return int(lib().axb_bruteforce_size(self._h))

    def empty(self):
        return self.size() == 0

    def bounds(self):
        out = (ctypes.c_float * 6)()
        _check(lib().axb_bruteforce_bounds(self._h, ctypes.cast(out, _P)))
        return [float(x) for x in out]

    def query(self, predicates, policy=None, return_distances=False, stream=None):
        """query(space, predicates, indices, offsets[, policy]) -> (indices, offsets)
        (ArborX_BruteForce.hpp:66-91); as BoundingVolumeHierarchy.query."""
        return _crs_query("axb_bruteforce", self._h, predicates, policy, return_distances,
                          stream)


BVH = BoundingVolumeHierarchy


def trim_memory():
    """Synchronises the device and returns the memory the library keeps cached
    (its stream-ordered pool and per-stream workspaces) to the driver."""
    _check(lib().axb_trim_memory())


def _neighbor_list(points, radius, full, stream):
    x = _dev_f32(points, 3)
    n = x.shape[0]
    offsets = torch.empty(n + 1, dtype=torch.int64, device="cuda")
    ind = ctypes.c_void_p()
    nnz = ctypes.c_int64()
    _use_torch_allocator()
    _check(lib().axb_find_neighbor_list(_ptr(x), n, float(radius), 1 if full else 0,
                                        _ptr(offsets), ctypes.byref(ind), ctypes.byref(nnz),
                                        _stream(stream)))
    return offsets, _take_output(ind.value, nnz.value, torch.int32)


def find_full_neighbor_list(points, radius, stream=None):
    """Experimental::findFullNeighborList(space, points, radius, offsets,
    indices) (detail/ArborX_NeighborList.hpp:105-188) -> (offsets, indices):
    row i lists every j != i within radius of point i."""
    return _neighbor_list(points, radius, True, stream)


def find_half_neighbor_list(points, radius, stream=None):
    """Experimental::findHalfNeighborList (detail/ArborX_NeighborList.hpp:49-103)
    -> (offsets, indices): each pair once, in the row of the point whose
    leaf comes later in the tree."""
    return _neighbor_list(points, radius, False, stream)


def expand_half_to_full(offsets, indices, stream=None):
    """Details::expandHalfToFull (detail/ArborX_ExpandHalfToFull.hpp:23-68)
    -> (offsets, indices) of the full list."""
    o = torch.as_tensor(offsets).to(device="cuda", dtype=torch.int64).contiguous()
    i = torch.as_tensor(indices).to(device="cuda", dtype=torch.int32).contiguous()
    n = o.shape[0] - 1
    out_o = torch.empty(n + 1, dtype=torch.int64, device="cuda")
    ind = ctypes.c_void_p()
    nnz = ctypes.c_int64()
    _use_torch_allocator()
    _check(lib().axb_expand_half_to_full(_ptr(o), _ptr(i) if i.numel() else None, n,
                                         _ptr(out_o), ctypes.byref(ind), ctypes.byref(nnz),
                                         _stream(stream)))
    return out_o, _take_output(ind.value, nnz.value, torch.int32)


findFullNeighborList = find_full_neighbor_list
findHalfNeighborList = find_half_neighbor_list
expandHalfToFull = expand_half_to_full


def query(tree, predicates, policy=None, **kw):
    """Free function ArborX::query(tree, space, predicates, indices, offsets)
    (src/spatial/ArborX_CrsGraphWrapper.hpp:22-34)."""
    return tree.query(predicates, policy, **kw)


def _cuda_memcpy_d2d(dst_tensor, src_ptr, nbytes):
    # Device-to-device copy of a raw pointer into a tensor via the library's
    # stream-ordered runtime (hipMemcpy through torch's loaded HIP runtime).
    hip = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    hip.hipMemcpy.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [_P, _P, ctypes.c_size_t, ctypes.c_int]
    err = hip.hipMemcpy(ctypes.c_void_p(dst_tensor.data_ptr()), ctypes.c_void_p(src_ptr),
                        nbytes, 3)  # hipMemcpyDeviceToDevice
    if err != 0:
        raise RuntimeError(f"hipMemcpy failed: {err}")


class profile:
    """Per-kernel timing and traversal visit counting of the library.

    with arborx_amd.profile(timing=True, visits=True) as p:
        ...                     # library calls
    p.kernels  -> {kernel name: [durations in ms]}
    p.visits   -> {"queries", "internal", "leaf", "results"}
    """

    def __init__(self, timing=True, visits=False):
        self.flags = (1 if timing else 0) | (2 if visits else 0)
        self.kernels = {}
        self.visits = None

    def __enter__(self):
        _check(lib().axb_profile_enable(self.flags))
        return self

    def __exit__(self, *exc):
        L = lib()
        torch.cuda.synchronize()
        buf = ctypes.create_string_buffer(256)
        for i in range(L.axb_profile_num_records()):
            ms = ctypes.c_double()
            _check(L.axb_profile_record(i, buf, 256, ctypes.byref(ms)))
            self.kernels.setdefault(buf.value.decode(), []).append(ms.value)
        if self.flags & 2:
            v = (_I64 * 12)()
            _check(L.axb_profile_visits(v))
            self.visits = dict(queries=v[0], internal=v[1], leaf=v[2], results=v[3],
                               fallback_queries=v[4], fallback_internal=v[5],
                               fallback_leaf=v[6], rounds=v[8], multi_round_queries=v[9])
        _check(L.axb_profile_enable(0))
        return False


class debug_limits:
    """Lowers limits the library derives from the device, so that small
    inputs take the paths of large ones (tests only; process-wide):

    with arborx_amd.debug_limits(knn_batch_max=1000, knn_rounds=1):
        ...   # kNN in launches of 1000 queries, most to the exact fallback
    """

    def __init__(self, knn_batch_max=0, scratch_rows_bytes=0, knn_rounds=0):
        self.args = (int(knn_batch_max), int(scratch_rows_bytes), int(knn_rounds))

    def __enter__(self):
        _check(lib().axb_debug_set_limits(*self.args))
        return self

    def __exit__(self, *exc):
        _check(lib().axb_debug_set_limits(0, 0, 0))
        return False


def q16_encode_host(frame6, box6):
    """Q16 steps {qlo xyz, qhi xyz} of a box in a frame (the build's encoder,
    run on the host)."""
    import numpy as np
    f = np.ascontiguousarray(frame6, np.float32)
    b = np.ascontiguousarray(box6, np.float32)
    q = np.zeros(6, np.uint16)
    lib().axb_q16_encode_host(f.ctypes.data_as(_P), b.ctypes.data_as(_P), q.ctypes.data_as(_P))
    return q


def sqrt_threshold(r):
    """Largest float s with RN(sqrt(s)) <= r (host helper of the library)."""
    return float(lib().axb_sqrt_threshold(ctypes.c_float(r)))


def morton64_host(p, scene):
    pa = (ctypes.c_float * 3)(*[float(x) for x in p])
    sa = (ctypes.c_float * 6)(*[float(x) for x in scene])
    return int(lib().axb_morton64_host(ctypes.cast(pa, _P), ctypes.cast(sa, _P)))


from .energy import PairEnergy, pair_energy, pair_energy_fd  # noqa: E402,F401
