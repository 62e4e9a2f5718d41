"""Writes tests/golden/reference_kat.json: known-answer vectors transcribed from
ArborX's own unit tests (file:line of each source test is stored with it).

The values are data copied from the tests' literals; where a test computes
its expectation with a formula (the Morton anchors of
tstDetailsTreeConstruction.cpp:66-72), the formula is evaluated here with an
independent bit-by-bit interleave, not with the oracle.
Run:  python tests/golden/make_fixtures.py
"""
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def interleave3(i, j, k, bits):
    """Closed-form 3-D Morton interleave: bit b of i -> bit 3b+2, j -> 3b+1,
    k -> 3b (the expansion the reference's anchor formula
    4*expandBitsBy<2>(i) + 2*expandBitsBy<2>(j) + expandBitsBy<2>(k) spells)."""
    r = 0
    for b in range(bits):
        r |= ((i >> b) & 1) << (3 * b + 2)
        r |= ((j >> b) & 1) << (3 * b + 1)
        r |= ((k >> b) & 1) << (3 * b)
    return r


def main():
    kat = {}
    # tstDetailsMortonCodes.cpp:22-49 (expandBitsBy<2> rows only: 3-D path)
    kat["expand_bits"] = {
        "source": "test/tstDetailsMortonCodes.cpp:22-49",
        "u32_in": 0b110010011101,
        "u32_expand2": 0b000000001000000001001001000001,
        "u64_in": 0b11111111111111000001,
        "u64_expand2": 0b001001001001001001001001001001001001001001000000000000000001,
    }
    # tstDetailsMortonCodes.cpp:51-80 (3-D rows)
    kat["morton_3d"] = {
        "source": "test/tstDetailsMortonCodes.cpp:71-79",
        "morton32": [
            [[0.0, 0.0, 0.0], 0x0],
            [[1.0, 1.0, 1.0], 0x3FFFFFFF],
            [[0.0, 0.0, 1.0], 0x9249249],
            [[1.0, 1.0, 0.0], 0x36DB6DB6],
        ],
        "morton64": [
            [[0.0, 0.0, 0.0], 0x0],
            [[1.0, 1.0, 1.0], 0x7FFFFFFFFFFFFFFF],
            [[0.0, 0.0, 1.0], 0x1249249249249249],
            [[1.0, 1.0, 0.0], 0x6DB6DB6DB6DB6DB6],
        ],
    }
    # tstDetailsTreeConstruction.cpp:33-102: 7 points as degenerate boxes,
    # scene box [0, 2^21]^3, codes = Morton of the anchors.
    N = 1 << 21
    pts = [[0.0, 0.0, 0.0], [0.25, 0.75, 0.25], [0.75, 0.25, 0.25], [0.75, 0.75, 0.25],
           [1.33, 2.33, 3.33], [1.66, 2.66, 3.66], [float(N)] * 3]
    anchors = [[0, 0, 0]] * 4 + [[1, 2, 3], [1, 2, 3], [N - 1] * 3]
    kat["assign_morton_codes"] = {
        "source": "test/tstDetailsTreeConstruction.cpp:33-102",
        "boxes": [[float(np.float32(c)) for c in p] * 2 for p in pts],
        "scene": [0.0, 0.0, 0.0, float(N), float(N), float(N)],
        "codes": [interleave3(a[0], a[1], a[2], 21) for a in anchors],
    }
    # tstDetailsTreeConstruction.cpp:119-149
    kat["indirect_sort"] = {
        "source": "test/tstDetailsTreeConstruction.cpp:119-149",
        "keys": [4, 3, 2, 1],
        "sorted": [1, 2, 3, 4],
        "perm": [3, 2, 1, 0],
    }
    # tstDetailsTreeConstruction.cpp:217-269 (Karras' example)
    kat["karras_8"] = {
        "source": "test/tstDetailsTreeConstruction.cpp:217-269",
        "codes": [int(b, 2) for b in ["00001", "00010", "00100", "00101",
                                      "10011", "11000", "11001", "11110"]],
        "dfs": "I0I3I1L0L1I2L2L3I4L4I5I6L5L6L7",
    }
    # tstQueryTreeDegenerate.cpp: trees of boxes, queries and CRS answers.
    # An empty query box is the default Box {+FLT_MAX, -FLT_MAX}.
    FM = float(np.finfo(np.float32).max)
    EMPTY = [FM, FM, FM, -FM, -FM, -FM]
    unit = [0.0, 0.0, 0.0, 1.0, 1.0, 1.0]
    p0 = [0.0] * 6
    p1 = [1.0] * 6
    kat["degenerate"] = {
        "source": "test/tstQueryTreeDegenerate.cpp:24-372",
        "cases": [
            # (tree boxes, predicate kind, predicates, indices, offsets)
            dict(line="133-150", tree=[unit], kind="boxes",
                 preds=[[5.0] * 6, [0.5] * 6], indices=[0], offsets=[0, 0, 1]),
            dict(line="152-165", tree=[unit], kind="spheres",
                 preds=[[0, 0, 0, 1.0], [1, 1, 1, 3.0], [5, 5, 5, 2.0]],
                 indices=[0, 0], offsets=[0, 1, 2, 2]),
            dict(line="210-216", tree=[unit], kind="nearest",
                 preds=[[0, 0, 0, 3], [4, 5, 1, 1]], indices=[0, 0], offsets=[0, 1, 2]),
            dict(line="218-224", tree=[unit], kind="nearest",
                 preds=[[0, 0, 0, 1], [1, 1, 1, 2], [2, 2, 2, 3]],
                 indices=[0, 0, 0], offsets=[0, 1, 2, 3]),
            dict(line="251-256", tree=[p0, p1], kind="boxes", preds=[EMPTY],
                 indices=[], offsets=[0, 0]),
            dict(line="258-263", tree=[p0, p1], kind="boxes", preds=[unit],
                 indices=[1, 0], offsets=[0, 2]),
            dict(line="265-270", tree=[p0, p1], kind="boxes",
                 preds=[[0.5] * 3 + [1.5] * 3], indices=[1], offsets=[0, 1]),
            dict(line="272-278", tree=[p0, p1], kind="boxes", preds=[EMPTY, EMPTY],
                 indices=[], offsets=[0, 0, 0]),
            dict(line="280-286", tree=[p0, p1], kind="boxes", preds=[EMPTY, p0],
                 indices=[0], offsets=[0, 0, 1]),
            dict(line="288-291", tree=[p0, p1], kind="boxes", preds=[],
                 indices=[], offsets=[0]),
            dict(line="322-328", tree=[p0, p1], kind="nearest",
                 preds=[[0, 0, 0, 2], [1, 0, 0, 4]], indices=[0, 1, 0, 1],
                 offsets=[0, 2, 4]),
            dict(line="350-367", tree=[p0, p1, p1, p1], kind="spheres",
                 preds=[[0, 0, 0, 1.0], [1, 1, 1, 1.0], [0.5, 0.5, 0.5, 1.0]],
                 indices=[0, 1, 2, 3, 0, 1, 2, 3], offsets=[0, 1, 4, 8]),
            dict(line="48-71", tree=[], kind="boxes", preds=[EMPTY, EMPTY],
                 indices=[], offsets=[0, 0, 0]),
            dict(line="69-71", tree=[], kind="spheres",
                 preds=[[0, 0, 0, 1.0], [1, 1, 1, 2.0]], indices=[], offsets=[0, 0, 0]),
            dict(line="117-121", tree=[], kind="nearest",
                 preds=[[0, 0, 0, 1], [1, 1, 1, 2]], indices=[], offsets=[0, 0, 0]),
        ],
        "note": "rows compare as multisets (CompressedStorageComparison.hpp:21-61)",
    }
    # tstQueryTreeDegenerate.cpp:375-446: 4096 diagonal unit boxes.
    kat["stack_capacity"] = {
        "source": "test/tstQueryTreeDegenerate.cpp:375-446",
        "n": 4096,
        "spatial_total": 4096,
        "nearest_total": 4096,
    }
    # tstQueryTreeCallbacks.cpp:138-179: points (i,i,i), i < 10; nearest(origin, 10)
    # returns every point; the callback's payload is distance to the origin.
    d = []
    for i in range(10):
        t = np.float32(i)
        s = np.float32(0)
        for _ in range(3):
            s = np.float32(s + t * t)
        d.append(float(np.sqrt(s, dtype=np.float32)))
    kat["nearest_collinear"] = {
        "source": "test/tstQueryTreeCallbacks.cpp:138-179,71-89",
        "points": [[float(i)] * 3 for i in range(10)],
        "query": [0.0, 0.0, 0.0],
        "k": 10,
        "indices": list(range(10)),
        "distances": d,
    }
    # tstQueryTreeTraversalPolicy.cpp buffer policy: setBufferSize(-1) with a
    # query having more than one result throws (:125-128).
    kat["buffer_policy"] = {
        "source": "test/tstQueryTreeTraversalPolicy.cpp:101-130",
        "note": "hard buffer overflow raises SearchException",
    }
    # tstNeighborList.cpp:116-176: full lists (also half lists expanded to
    # full) of four collinear points; rows compare as multisets.  radius
    # "inf" stands for the test's infinity<float>.
    kat["neighbor_list"] = {
        "source": "test/tstNeighborList.cpp:116-176",
        "degenerate": [
            {"points": [], "radius": "inf", "offsets": [0], "indices": []},
            {"points": [[0.0, 0.0, 0.0]], "radius": "inf", "offsets": [0, 0], "indices": []},
        ],
        "points": [[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [2.0, 2.0, 2.0], [3.0, 3.0, 3.0]],
        "cases": [
            {"radius": 1.0, "offsets": [0, 0, 0, 0, 0], "indices": []},
            {"radius": 2.0, "offsets": [0, 1, 3, 5, 6], "indices": [1, 0, 2, 1, 3, 2]},
            {"radius": 4.0, "offsets": [0, 2, 5, 8, 10],
             "indices": [1, 2, 0, 2, 3, 0, 1, 3, 1, 2]},
            {"radius": 6.0, "offsets": [0, 3, 6, 9, 12],
             "indices": [1, 2, 3, 0, 2, 3, 0, 1, 3, 0, 1, 2]},
        ],
    }
    # tstQueryTreeRay.cpp: ten unit boxes [i, i+1]^3 on the diagonal and rays
    # given as origin + (unnormalised) direction; "ordered" rows are compared
    # as sequences, the others as sets.  The single-leaf case is
    # :212-243 (one box [0,1]^3), the empty tree :185-210.
    diag = [[float(i)] * 3 + [float(i + 1)] * 3 for i in range(10)]
    kat["rays"] = {
        "source": "test/tstQueryTreeRay.cpp:29-243",
        "cases": [
            {"line": "29-58", "boxes": diag, "kind": "nearest", "k": 1,
             "origins": [[0.0, 0.0, 0.0]], "directions": [[0.15, 0.1, 0.0]],
             "offsets": [0, 1], "indices": [0]},
            {"line": "60-87", "boxes": diag, "kind": "intersects",
             "origins": [[0.0, 0.0, 0.0]], "directions": [[0.1, 0.1, 0.1]],
             "offsets": [0, 10], "indices": list(range(10))},
            {"line": "112-163", "boxes": diag, "kind": "ordered",
             "origins": [[0.0, 0.0, 0.0], [10.0, 10.0, 10.0]],
             "directions": [[0.1, 0.1, 0.1], [-0.1, -0.1, -0.1]],
             "offsets": [0, 10, 20], "indices": list(range(10)) + list(range(9, -1, -1))},
            {"line": "185-210", "boxes": [], "kind": "ordered",
             "origins": [[0.0, 0.0, 0.0], [0.0, 0.0, 0.0]],
             "directions": [[1.0, 0.0, 0.0], [1.0, 0.0, 0.0]],
             "offsets": [0, 0, 0], "indices": []},
            {"line": "212-243", "boxes": [[0.0, 0.0, 0.0, 1.0, 1.0, 1.0]], "kind": "ordered",
             "origins": [[0.0, 0.0, 0.0], [4.0, 5.0, 6.0]],
             "directions": [[1.0, 1.0, 1.0], [7.0, 8.0, 9.0]],
             "offsets": [0, 1, 1], "indices": [0]},
        ],
    }
    # tstQueryTreeCallbackQueryPerThread.cpp:81-110: ten points (i, i, i),
    # attach_indices values; query i is ordered_intersects of the box
    # [i - .5, i + .5]^3 with a callback that records whether the first
    # value it sees is i, then returns early_exit.  (:36-79: the same with
    # intersects(box), every match being i.)
    pts10 = [[float(i)] * 3 for i in range(10)]
    boxes10 = [[i - 0.5] * 3 + [i + 0.5] * 3 for i in range(10)]
    kat["per_thread"] = {
        "source": "test/tstQueryTreeCallbackQueryPerThread.cpp:36-110",
        "points": pts10, "query_boxes": boxes10, "first_value": list(range(10)),
    }
    with open(os.path.join(HERE, "reference_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)
    print("wrote", os.path.join(HERE, "reference_kat.json"))


if __name__ == "__main__":
    main()
