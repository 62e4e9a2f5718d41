"""Runs tests/cpp/test_walk_host (built by __graft_entry__.build()): the
library's traversal code from include/arborx_mi355x_device.hpp -- the
function the GPU kernels run -- executed on the CPU over node arrays written
by the oracle, against the oracle's reference-form rope traversal (rows
identical in content and order; FP32 / Q16 nodes, flattened subtrees, leaf
groups, the neighbour lists' skipped leaf and shortened walk).  CPU only."""
import os
import subprocess

from tests.conftest import ROOT

BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_walk_host")


def test_device_walk_on_host_matches_oracle():
    assert os.path.exists(BIN), f"{BIN} missing: run __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert " 0 failed" in r.stdout
