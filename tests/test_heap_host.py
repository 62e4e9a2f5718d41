"""Runs tests/cpp/test_heap_host (built by __graft_entry__.build()): the
(distance, index) max-heap of the kNN output rows for k > 64, from
include/arborx_mi355x_device.hpp, on the CPU -- the reference's heap
insertion sequences (test/tstHeapOperations.cpp) and top-k against a plain
partial sort.  CPU only."""
import os
import subprocess

from tests.conftest import ROOT

BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_heap_host")


def test_row_heap_on_host():
    assert os.path.exists(BIN), f"{BIN} missing: run __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert " 0 failed" in r.stdout
