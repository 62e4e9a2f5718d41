"""Runs the C++ facade test program (tests/cpp/test_facade.cpp, built by
__graft_entry__.build()) on the GPU: the reference-shaped C++ API over the C
ABI, checked with the reference's degenerate / callback / manufactured /
policy cases and brute force."""
import os
import subprocess

import pytest

from tests.conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_facade")


def test_cpp_facade_program():
    assert os.path.exists(BIN), "tests/cpp/_build/test_facade missing: run __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert " 0 failed" in r.stdout
