"""CPU-side checks of the C-ABI library: it loads, exports every symbol that
include/*.h declares, and its host helpers are exact.  No GPU needed."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import arborx_amd
from tests.conftest import ROOT


def _declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//.*", "", src)
        names |= set(re.findall(r"\b(axb_\w+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(arborx_amd.LIB_PATH)
    declared = _declared_symbols()
    assert len(declared) >= 18
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    assert set(arborx_amd.EXPORTED_SYMBOLS) == declared


def test_library_marks_ranges_with_roctx():
    """The entry points open roctx ranges (the reference's profiling
    region labels): the library imports roctx's push and pop and holds the
    labels of the build phases and the query forms."""
    import subprocess
    und = subprocess.run(["nm", "-D", "--undefined-only", arborx_amd.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    assert "roctxRangePushA" in und and "roctxRangePop" in und
    blob = open(arborx_amd.LIB_PATH, "rb").read()
    for label in (b"ArborX::BVH::BVH", b"ArborX::BVH::BVH::sort_linearized_order",
                  b"ArborX::BVH::BVH::generate_hierarchy", b"ArborX::BVH::query_crs",
                  b"ArborX::BVH::query::spatial", b"ArborX::BVH::query::nearest",
                  b"ArborX::CrsGraphWrapper::two_pass",
                  b"ArborX::Experimental::FullNeighborList",
                  b"ArborX::BruteForce::BruteForce"):
        assert label + b"\0" in blob, label


def test_version_string():
    assert b"gfx950" in arborx_amd.lib().axb_version()


def test_sqrt_threshold_is_exact():
    """d2 <= T(r) must decide exactly RN(sqrt(d2)) <= r (numpy's float32 sqrt is
    correctly rounded)."""
    rng = np.random.default_rng(0)
    radii = np.concatenate([
        rng.uniform(1e-4, 10, 3000).astype(np.float32),
        np.float32([0.0336778, 0.00914156, 1.0, 0.5, 2.0, 1e-30, 3e18, 1e-3]),
    ])
    for r in radii:
        t = np.float32(arborx_amd.sqrt_threshold(float(r)))
        assert np.sqrt(t, dtype=np.float32) <= r
        nxt = np.nextafter(t, np.float32(np.inf))
        assert np.sqrt(nxt, dtype=np.float32) > r
        # random d2 near r^2
        base = np.float32(r) * np.float32(r)
        cand = base + (rng.integers(-40, 40, 64).astype(np.float32) * np.spacing(base))
        cand = cand[cand >= 0].astype(np.float32)
        np.testing.assert_array_equal(np.sqrt(cand, dtype=np.float32) <= r, cand <= t)


def test_sqrt_threshold_special_values():
    assert arborx_amd.sqrt_threshold(0.0) == 0.0
    assert arborx_amd.sqrt_threshold(float("inf")) == float("inf")
    assert arborx_amd.sqrt_threshold(-1.0) < 0
    fm = float(np.finfo(np.float32).max)
    assert arborx_amd.sqrt_threshold(fm) == fm


def test_morton64_host_matches_oracle(kat):
    import oracle
    for p, code in kat["morton_3d"]["morton64"]:
        assert arborx_amd.morton64_host(p, [0, 0, 0, 1, 1, 1]) == code
    from arborx_amd import synthetic
    pts = synthetic.uniform_points(500, 3)
    scene = oracle.scene_box(pts)
    ref = oracle.morton64_values(pts, scene)
    for p, c in zip(pts, ref):
        assert arborx_amd.morton64_host(p, scene) == int(c)


def test_policy_and_predicates_mirror_reference_defaults():
    p = arborx_amd.TraversalPolicy()
    assert p._buffer_size == 0 and p._sort_predicates is True  # TraversalPolicy.hpp:31,34
    assert p.setBufferSize(-1).setPredicateSorting(False)._c().buffer_size == -1
    with pytest.raises(ValueError):
        arborx_amd.make_intersects(np.zeros((1, 3)), 0.0)  # PredicateHelpers.hpp:102


def test_flat_subtree_constant_matches_header():
    import re
    with open(os.path.join(ROOT, "include", "arborx_mi355x.h")) as f:
        m = re.search(r"#define AXB_FLAT_SUBTREE_MAX (\d+)", f.read())
    assert m and int(m.group(1)) == arborx_amd.FLAT_SUBTREE_MAX


def test_q16_encoder_matches_oracle():
    """The build's Q16 encoder (bracketed search from an estimate) against the
    oracle's plain binary search, on random, tiny, huge and degenerate
    frames and on boxes touching the frame."""
    import oracle
    rng = np.random.default_rng(11)
    for t in range(4000):
        scale = [1e-30, 1e-6, 1.0, 1e6, 1e30][t % 5]
        f = np.sort(rng.uniform(-1, 1, (2, 3)) * scale, axis=0).astype(np.float32)
        if t % 97 == 0:
            f[1, 0] = f[0, 0]  # flat axis
        frame = np.concatenate([f[0], f[1]])
        u = np.sort(rng.uniform(0, 1, (2, 3)), axis=0)
        if t % 7 == 0:
            u[0, :] = 0.0
            u[1, :] = 1.0
        box = np.concatenate([f[0] + (f[1] - f[0]) * u[0], f[0] + (f[1] - f[0]) * u[1]])
        box = box.astype(np.float32)
        box = np.concatenate([np.maximum(box[:3], frame[:3]), np.minimum(box[3:], frame[3:])])
        q_ref, _ = oracle.q16_encode_box(frame, box)
        np.testing.assert_array_equal(arborx_amd.q16_encode_host(frame, box), q_ref, err_msg=str(t))
