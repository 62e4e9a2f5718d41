"""axb_exclusive_scan_i64 (the CRS offsets' scan, KokkosExt::exclusive_scan)
against numpy's cumulative sum: tile edges (2048 elements per tile), more
tiles than one chunk of the tile-offset pass (8192 tiles), in place and out
of place, and values whose sum needs all 64 bits' range below 2^62."""
import ctypes

import numpy as np
import pytest
import torch

import arborx_amd as ax

pytestmark = pytest.mark.gpu


def _scan(x, out):
    f = ax.lib().axb_exclusive_scan_i64
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    f.restype = ctypes.c_int32
    st = f(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()), x.numel(),
           ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == 0, ax.lib().axb_last_error()
    torch.cuda.synchronize()


def _want(h):
    return np.concatenate([[0], np.cumsum(h)[:-1]]).astype(np.int64)


@pytest.mark.parametrize("n", [1, 2, 7, 2047, 2048, 2049, 4097, 100_003,
                               8192 * 2048, 8192 * 2048 + 1, 40_000_000])
def test_exclusive_scan_matches_numpy(n):
    print(f"scan n={n}", flush=True)
    rng = np.random.default_rng(n)
    h = rng.integers(0, 1000, size=n, dtype=np.int64)
    h[rng.integers(0, n, size=max(1, n // 1000))] = 1 << 40  # large entries
    want = _want(h)
    x = torch.from_numpy(h).cuda()
    out = torch.full_like(x, -1)
    _scan(x, out)
    np.testing.assert_array_equal(out.cpu().numpy(), want)
    # in place, and the input of the out-of-place call left untouched
    np.testing.assert_array_equal(x.cpu().numpy(), h)
    _scan(x, x)
    np.testing.assert_array_equal(x.cpu().numpy(), want)


def test_exclusive_scan_empty_is_a_no_op():
    x = torch.zeros(0, dtype=torch.int64, device="cuda")
    _scan(x, x)
