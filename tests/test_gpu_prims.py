"""gpu-prims on the device: scan, stable radix sort, reduce, Threefry, 64-bit atomics and the
kick conversion, checked against numpy / the oracle at the reference's edge sizes
(crates/gpu-prims/src/checks.rs:31)."""
import math

import numpy as np
import pytest

from micrographia_amd import _lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 255, 256, 257, 1023, 1024, 1025, 4097, 65537, (1 << 20) + 3]


def torch_cuda():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def dptr(t):
    return t.data_ptr()


@pytest.mark.parametrize("n", SIZES)
def test_exclusive_scan(n):
    torch = torch_cuda()
    rng = np.random.default_rng(n)
    a = rng.integers(0, 1000, n, dtype=np.uint32)
    x = torch.from_numpy(a.view(np.int32)).cuda()
    y = torch.empty_like(x)
    _lib.check(_lib.load().rh_prims_exclusive_scan_u32(dptr(x), dptr(y), n, None))
    csum = np.cumsum(a.astype(np.uint64))
    want = (np.concatenate([np.zeros(1, np.uint64), csum[:-1]]) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    assert np.array_equal(y.cpu().numpy().view(np.uint32), want)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("bits", [8, 18, 32])
def test_radix_sort_is_stable(n, bits):
    torch = torch_cuda()
    rng = np.random.default_rng(n + bits)
    hi = 1 << bits if bits < 32 else 1 << 32
    keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    k = torch.from_numpy(keys.view(np.int32)).cuda()
    v = torch.from_numpy(vals.view(np.int32)).cuda()
    _lib.check(_lib.load().rh_prims_radix_sort_u32(dptr(k), dptr(v), n, bits, None))
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k.cpu().numpy().view(np.uint32), keys[order])
    assert np.array_equal(v.cpu().numpy().view(np.uint32), vals[order])


def test_radix_sort_one_key_adversarial():
    """Every item in one bucket (gpu-prims checks.rs one-key bin)."""
    torch = torch_cuda()
    n = 200003
    keys = np.full(n, 77, np.uint32)
    k = torch.from_numpy(keys.view(np.int32)).cuda()
    v = torch.arange(n, dtype=torch.int32).cuda()
    _lib.check(_lib.load().rh_prims_radix_sort_u32(dptr(k), dptr(v), n, 18, None))
    assert np.array_equal(v.cpu().numpy(), np.arange(n))


@pytest.mark.parametrize("n", [1, 2, 257, 2049, 1 << 20])
def test_reduce_sum_f64(n):
    torch = torch_cuda()
    a = np.random.default_rng(n).random(n)
    x = torch.from_numpy(a).cuda()
    out = torch.zeros(1, dtype=torch.float64).cuda()
    _lib.check(_lib.load().rh_prims_reduce_sum_f64(dptr(x), dptr(out), n, None))
    assert abs(out.item() - a.sum()) <= 1e-12 * n


def test_threefry_matches_the_cpu_twin():
    """65,536 draws on the device against the oracle (gpu-prims checks.rs)."""
    torch = torch_cuda()
    n = 65536
    out = torch.zeros(2 * n, dtype=torch.int32).cuda()
    _lib.check(_lib.load().rh_prims_threefry(dptr(out), n, 11, 3, 0xDEADBEEF, 0x12345678, None))
    got = out.cpu().numpy().view(np.uint32).reshape(n, 2)
    for i in list(range(0, n, 997)) + [n - 1]:
        assert tuple(int(x) for x in got[i]) == O.threefry((i + 11, 3), (0xDEADBEEF, 0x12345678))


def test_atomic_add_i64():
    """200,003 64-bit atomic adds (gpu-prims checks.rs), with carries across the low word."""
    torch = torch_cuda()
    n = 200003
    vals = np.random.default_rng(3).integers(-(1 << 40), 1 << 40, n, dtype=np.int64)
    x = torch.from_numpy(vals).cuda()
    acc = torch.zeros(1, dtype=torch.int64).cuda()
    _lib.check(_lib.load().rh_prims_atomic_add_i64(dptr(x), n, dptr(acc), None))
    assert acc.item() == int(vals.sum())


def rust_kick(x: float) -> int:
    """round_ties_even then Rust's saturating `as i32` (NaN -> 0)."""
    if math.isnan(x):
        return 0
    r = float(np.rint(np.float32(x)))
    if r >= 2147483647.0:
        return 2147483647
    if r <= -2147483648.0:
        return -2147483648
    return int(r)


def test_kick_conversion_on_the_hardware():
    torch = torch_cuda()
    specials = [0.0, -0.0, 0.5, 1.5, 2.5, -0.5, -1.5, -2.5, 3.4999998, 1e9, -1e9, 2147483520.0, 2147483648.0, 3e9, -3e9,
                -2147483648.0, float("inf"), float("-inf"), float("nan"), 1e-40, -7.5, 8388607.5, 8388608.5]
    rng = np.random.default_rng(1)
    rand = (rng.standard_normal(100000) * 10.0 ** rng.integers(-2, 11, 100000)).astype(np.float32)
    xs = np.concatenate([np.array(specials, np.float32), rand])
    x = torch.from_numpy(xs).cuda()
    out = torch.zeros(len(xs), dtype=torch.int32).cuda()
    _lib.check(_lib.load().rh_prims_kick_convert(dptr(x), dptr(out), len(xs), None))
    got = out.cpu().numpy()
    want = np.array([rust_kick(float(v)) for v in xs], np.int64)
    assert np.array_equal(got.astype(np.int64), want)


@pytest.mark.parametrize("eps,sigma,rc", [(1.0, 1.0, 2.5), (1.0, 1.0, 2.0 ** (1 / 6)), (0.5, 1.2, 3.0), (0.7, 1.05, 2.5), (1.0, 0.9, 2.2)])
def test_shared_reciprocal_division_is_ieee_for_every_r2(eps, sigma, rc):
    """Exhaustive: for every f32 r2 the fast path handles, the pair's two divisions equal IEEE
    division bit for bit (so the kicks equal md-ref's)."""
    torch_cuda()
    import ctypes

    checked, mism = ctypes.c_uint64(), ctypes.c_uint64()
    first = ctypes.c_float()
    _lib.check(_lib.load().rh_prims_check_pair_division(eps, sigma, rc, ctypes.byref(checked), ctypes.byref(mism), ctypes.byref(first), None))
    assert checked.value > 10_000_000
    assert mism.value == 0, f"first differing r2 {first.value!r}"


def test_stochastic_draws_match_the_host_c_library():
    """The Andersen bath's and the heat-bath walls' draws (three.rs:232-265) on 100,000,000
    (id, clock) keys: the device's f64 log / cos / sin against the host C library's (which
    md-ref's f64 methods call), compared as the integer displacements the leapfrog stores.  A
    one-ulp difference matters only where it straddles a rounding boundary of that
    displacement; the test counts those."""
    import concurrent.futures as cf

    torch = torch_cuda()
    lib = _lib.load()
    seed, t, m, dt = 0x5EED_0000_0000_0007, 1.3, 1.0, 0.005
    chunk = 10_000_000
    out = torch.empty(chunk * 6, dtype=torch.int32, device="cuda")
    total = mismatched_keys = mismatched_values = 0
    for c in range(10):
        id0, clock = c * chunk, 1000 + 7 * c
        print(f"draws: chunk {c}, keys {id0}..{id0 + chunk}", flush=True)
        _lib.check(lib.rh_prims_stochastic_draws(seed, t, m, dt, id0, chunk, clock, dptr(out), None))
        got = out.cpu().numpy().reshape(chunk, 6)
        parts = 16
        step = chunk // parts
        with cf.ThreadPoolExecutor(max_workers=parts) as ex:
            want = np.concatenate(list(ex.map(lambda p: O.draws3(seed, t, m, dt, id0 + p * step, step, clock), range(parts))))
        diff = got != want
        total += chunk
        mismatched_keys += int(diff.any(axis=1).sum())
        mismatched_values += int(diff.sum())
        if diff.any():
            k = int(np.argmax(diff.any(axis=1)))
            print(f"  first mismatch at key {id0 + k}: device {got[k]} host {want[k]}", flush=True)
    rate = mismatched_keys / total
    print(f"draws: {total} keys x 6 values, {mismatched_keys} keys ({mismatched_values} values) differ: rate {rate:.3g} per key")
    assert rate < 1e-6
