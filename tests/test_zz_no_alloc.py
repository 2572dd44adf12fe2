"""The radix sort that allocates nothing (rh_prims_radix_sort_u32_noalloc) against the allocating
entry point (rh_prims_radix_sort_u32), through the C ABI: two calls in a row on the same GPU, the
second with fewer items (so it runs over what the first left in the library's fixed scratch), at
odd and even pass counts (the result lands in the scratch half of the ping-pong pair, or in the
caller's), plus the size limit."""
import ctypes

import numpy as np
import pytest

from micrographia_amd import _lib

pytestmark = pytest.mark.gpu

MAX = 1 << 20  # RH_SORT_NOALLOC_MAX


def _torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def _sort_both(lib, keys, vals, bits):
    """Sort (keys, vals) with both entry points; returns both results and numpy's stable sort."""
    torch = _torch()
    ka = torch.from_numpy(keys.view(np.int32)).cuda()
    va = torch.from_numpy(vals.view(np.int32)).cuda()
    kb, vb = ka.clone(), va.clone()
    n = keys.size
    _lib.check(lib.rh_prims_radix_sort_u32(ka.data_ptr(), va.data_ptr(), n, bits, None))
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.rh_prims_radix_sort_u32_noalloc(kb.data_ptr(), vb.data_ptr(), n, bits, ctypes.c_void_p(stream)))
    torch.cuda.synchronize()
    return ((ka.cpu().numpy().view(np.uint32), va.cpu().numpy().view(np.uint32)),
            (kb.cpu().numpy().view(np.uint32), vb.cpu().numpy().view(np.uint32)))


@pytest.mark.parametrize("bits", [8, 18, 24, 32])
def test_two_calls_in_a_row_match_the_allocating_sort(bits):
    lib = _lib.load()
    rng = np.random.default_rng(bits)
    for n in (MAX, 65537, 1):
        print(f"no-alloc sort: n={n} bits={bits}", flush=True)
        hi = 1 << bits
        keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
        vals = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        (ka, va), (kb, vb) = _sort_both(lib, keys, vals, bits)
        assert np.array_equal(kb, ka) and np.array_equal(vb, va), (n, bits)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(kb, keys[order]) and np.array_equal(vb, vals[order]), (n, bits)


def test_many_equal_keys_keep_their_order():
    lib = _lib.load()
    n = 300001
    keys = np.full(n, 7, np.uint32)
    keys[::97] = 3
    vals = np.arange(n, dtype=np.uint32)
    (ka, va), (kb, vb) = _sort_both(lib, keys, vals, 16)
    assert np.array_equal(kb, ka) and np.array_equal(vb, va)


def test_sizes_beyond_the_scratch_are_refused():
    torch = _torch()
    lib = _lib.load()
    k = torch.arange(4, dtype=torch.int32, device="cuda")
    before = k.clone()
    rc = lib.rh_prims_radix_sort_u32_noalloc(k.data_ptr(), k.data_ptr(), MAX + 1, 32, None)
    assert rc == -1 and b"exceeds" in lib.rh_last_error()
    assert lib.rh_prims_radix_sort_u32_noalloc(k.data_ptr(), k.data_ptr(), 4, 0, None) == -1
    assert lib.rh_prims_radix_sort_u32_noalloc(None, None, 0, 32, None) == 0
    torch.cuda.synchronize()
    assert torch.equal(k, before)
