"""axb_pair_energy_noalloc (the allocation-free form of axb_pair_energy, its
partial sums in a static device array) against axb_pair_energy, both called
through the C ABI: energy and gradient bit-identical on two calls in a row,
the second with a smaller input, plus the size limit and the argument
checks.  Inputs below the size at which axb_pair_energy reorders its rows,
so both sum the same partials in the same order."""
import ctypes

import numpy as np
import pytest
import torch

import arborx_amd as ax
from arborx_amd import synthetic

pytestmark = pytest.mark.gpu

MAX_N = 1 << 27  # AXB_PAIR_ENERGY_NOALLOC_MAX_N


def _call(fn, x, off, idx, rc):
    n = x.shape[0]
    e = torch.full((), -1.0, dtype=torch.float64, device="cuda")
    g = torch.full((n, 3), float("nan"), dtype=torch.float32, device="cuda")
    st = fn(ax._ptr(x), n, ax._ptr(off), ax._ptr(idx), ctypes.c_float(rc),
            ctypes.c_void_p(e.data_ptr()), ax._ptr(g), ax._stream())
    assert st == 0, ax.lib().axb_last_error()
    torch.cuda.synchronize()
    return e.item(), g.cpu().numpy()


@pytest.mark.parametrize("cells", [(24, 24, 20), (9, 7, 5)])
def test_noalloc_energy_equals_allocating_entry_point(cells):
    lib = ax.lib()
    for c in (cells, (cells[0] // 2 + 1, cells[1] // 2 + 1, cells[2] // 2 + 1)):
        x = torch.from_numpy(synthetic.fcc_lattice(c, seed=6)).cuda()
        rc = 2.5
        off, idx = ax.energy.neighbour_csr(x, rc)
        e_ref, g_ref = _call(lib.axb_pair_energy, x, off, idx, rc)
        e_new, g_new = _call(lib.axb_pair_energy_noalloc, x, off, idx, rc)
        assert np.isfinite(e_ref) and e_ref < 0
        assert np.float64(e_new).view(np.uint64) == np.float64(e_ref).view(np.uint64)
        np.testing.assert_array_equal(g_new.view(np.uint32), g_ref.view(np.uint32))


def test_noalloc_energy_only_and_gradient_only():
    lib = ax.lib()
    x = torch.from_numpy(synthetic.fcc_lattice((6, 6, 6), seed=6)).cuda()
    off, idx = ax.energy.neighbour_csr(x, 2.5)
    e_ref, g_ref = _call(lib.axb_pair_energy, x, off, idx, 2.5)
    n = x.shape[0]
    e = torch.zeros((), dtype=torch.float64, device="cuda")
    assert lib.axb_pair_energy_noalloc(ax._ptr(x), n, ax._ptr(off), ax._ptr(idx),
                                       ctypes.c_float(2.5), ctypes.c_void_p(e.data_ptr()),
                                       None, ax._stream()) == 0
    g = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    assert lib.axb_pair_energy_noalloc(ax._ptr(x), n, ax._ptr(off), ax._ptr(idx),
                                       ctypes.c_float(2.5), None, ax._ptr(g),
                                       ax._stream()) == 0
    torch.cuda.synchronize()
    assert e.item() == e_ref
    np.testing.assert_array_equal(g.cpu().numpy().view(np.uint32), g_ref.view(np.uint32))


def test_noalloc_energy_rejects_oversized_and_bad_arguments():
    lib = ax.lib()
    x = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    off = torch.zeros(5, dtype=torch.int64, device="cuda")
    assert lib.axb_pair_energy_noalloc(ax._ptr(x), MAX_N + 1, ax._ptr(off), None,
                                       ctypes.c_float(2.5), None, None, ax._stream()) != 0
    assert b"AXB_PAIR_ENERGY_NOALLOC_MAX_N" in lib.axb_last_error()
    assert lib.axb_pair_energy_noalloc(ax._ptr(x), 4, ax._ptr(off), None,
                                       ctypes.c_float(0.0), None, None, ax._stream()) != 0
    assert lib.axb_pair_energy_noalloc(None, 4, None, None, ctypes.c_float(2.5), None, None,
                                       ax._stream()) != 0
    # an empty set is fine: E = 0
    e = torch.full((), 7.0, dtype=torch.float64, device="cuda")
    assert lib.axb_pair_energy_noalloc(None, 0, None, None, ctypes.c_float(2.5),
                                       ctypes.c_void_p(e.data_ptr()), None, ax._stream()) == 0
    torch.cuda.synchronize()
    assert e.item() == 0.0
