"""The scan entry point that allocates nothing (w2w_exclusive_scan_u32_noalloc)
against the allocating one (w2w_exclusive_scan_u32), through the C ABI: two
calls in a row, the second smaller, so that the shared static scratch is
reused and must be cleared; results bit-identical (integers)."""
import numpy as np
import pytest
import torch

from word2world_amd import _lib

pytestmark = pytest.mark.gpu

MAX_N = 1 << 28  # W2W_SCAN_NOALLOC_MAX_N


def _scan_alloc(L, x, s):
    n = x.numel()
    out = torch.empty(max(n, 1), dtype=torch.int32, device=x.device)
    tot = torch.zeros(1, dtype=torch.int32, device=x.device)
    tb = L.w2w_scan_temp_bytes(n)
    temp = torch.empty(max(tb, 1), dtype=torch.uint8, device=x.device)
    _lib.check(L.w2w_exclusive_scan_u32(x.data_ptr(), out.data_ptr(), tot.data_ptr(), n, temp.data_ptr(),
                                        tb, s), "w2w_exclusive_scan_u32")
    return out[:n], tot


def _scan_noalloc(L, x, s):
    n = x.numel()
    out = torch.full((max(n, 1),), -1, dtype=torch.int32, device=x.device)
    tot = torch.full((1,), -1, dtype=torch.int32, device=x.device)
    _lib.check(L.w2w_exclusive_scan_u32_noalloc(x.data_ptr(), out.data_ptr(), tot.data_ptr(), n, s),
               "w2w_exclusive_scan_u32_noalloc")
    return out[:n], tot


def test_noalloc_scan_matches_allocating_scan(gpu):
    L = _lib.load()
    s = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(17)
    for n in (6_000_011, 100_003, 16384, 1):  # each call smaller than the one before
        print(f"noalloc scan n={n}", flush=True)
        x = torch.from_numpy(rng.integers(0, 160, n).astype(np.int32)).to(gpu)  # totals below 2^30
        ref, ref_tot = _scan_alloc(L, x, s)
        got, got_tot = _scan_noalloc(L, x, s)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), n
        assert int(got_tot[0]) == int(ref_tot[0]) == int(x.sum())
    # n = 0 writes a zero total and touches nothing else
    tot = torch.full((1,), -1, dtype=torch.int32, device=gpu)
    _lib.check(L.w2w_exclusive_scan_u32_noalloc(None, None, tot.data_ptr(), 0, s), "n=0")
    torch.cuda.synchronize()
    assert int(tot[0]) == 0
    # beyond the stated maximum: refused before any launch
    dummy = torch.zeros(4, dtype=torch.int32, device=gpu)
    assert L.w2w_exclusive_scan_u32_noalloc(dummy.data_ptr(), dummy.data_ptr(), dummy.data_ptr(),
                                            MAX_N + 1, s) == 3
