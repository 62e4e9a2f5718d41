"""Runs the C++ facade test programs (tests/cpp/, built by
__graft_entry__.build()) on the GPU: test_facade -- the reference-shaped C++
API over the C ABI, checked with the reference's degenerate / callback /
manufactured / policy / neighbour-list cases and brute force; test_callbacks
-- the header-only callback traversals (tstQueryTreeCallbacks.cpp cases,
early exit, query(PerThread) in a user kernel)."""
import os
import subprocess

import pytest

from tests.conftest import ROOT

pytestmark = pytest.mark.gpu

BUILD = os.path.join(ROOT, "tests", "cpp", "_build")


@pytest.mark.parametrize("program", ["test_facade", "test_callbacks"])
def test_cpp_program(program):
    path = os.path.join(BUILD, program)
    assert os.path.exists(path), f"{path} missing: run __graft_entry__.build()"
    print(f"running {program}", flush=True)
    r = subprocess.run([path], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert " 0 failed" in r.stdout
