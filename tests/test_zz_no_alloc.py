"""mpm_sort_keys_noalloc: the engine's stable key sort on device buffers, with its temporary buffers
in the library's fixed device array instead of allocated memory.  Checked against the allocating
entry point (mpm_sort_keys) and numpy's stable argsort, bit for bit, on two calls in a row (the
second smaller, so the array keeps the first call's contents), and its size limit."""
import numpy as np
import pytest
import torch

from nuclearmpm_amd._lib import check, lib, ptr

pytestmark = pytest.mark.gpu


def host_sort(keys, bits):
    perm = np.empty(len(keys), np.uint32)
    check(lib().mpm_sort_keys(ptr(keys), len(keys), bits, ptr(perm), 0))
    return perm


def test_noalloc_sort_equals_the_allocating_sort_on_two_calls():
    rng = np.random.default_rng(3)
    # 3 radix passes (17 bits) on the largest size, then 2 passes on a smaller one
    for n, bits in ((1 << 20, 17), (70001, 12)):
        keys = rng.integers(0, 1 << bits, n, dtype=np.uint32)
        keys[1000:5000] = keys[1000]  # a long run of equal keys (stability)
        dk = torch.from_numpy(keys.view(np.int32)).cuda()
        dp = torch.full((n,), -1, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        print(f"[noalloc] n={n} bits={bits}", flush=True)
        check(lib().mpm_sort_keys_noalloc(dk.data_ptr(), n, bits, dp.data_ptr(), None, 0))
        torch.cuda.synchronize()
        got = dp.cpu().numpy().view(np.uint32)
        np.testing.assert_array_equal(got, host_sort(keys, bits))
        np.testing.assert_array_equal(got, np.argsort(keys, kind="stable").astype(np.uint32))
        np.testing.assert_array_equal(dk.cpu().numpy().view(np.uint32), keys)  # input unchanged


def test_noalloc_sort_limits_and_empty_input():
    L = lib()
    dk = torch.zeros(16, dtype=torch.int32, device="cuda")
    assert L.mpm_sort_keys_noalloc(dk.data_ptr(), (1 << 20) + 1, 8, dk.data_ptr(), None, 0) == 1
    assert b"at most 1048576 keys" in L.mpm_last_error()
    assert L.mpm_sort_keys_noalloc(None, 0, 8, None, None, 0) == 0
    assert L.mpm_sort_keys_noalloc(dk.data_ptr(), 4, 33, dk.data_ptr(), None, 0) == 1
