"""Runs tests/cpp/test_select_host (built by __graft_entry__.build()): the
kNN collect kernel's row selection from include/arborx_mi355x_device.hpp
executed on the CPU against a plain sort by (RN(sqrt(d2)), index), on rows
with ulp-close and repeated squared distances.  CPU only."""
import os
import subprocess

from tests.conftest import ROOT

BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_select_host")


def test_row_selection_on_host_matches_sort():
    assert os.path.exists(BIN), f"{BIN} missing: run __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert " 0 failed" in r.stdout
