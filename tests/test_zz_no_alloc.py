"""msplat_hip_radix_sort_noalloc (the entry point that allocates nothing; its scratch is a
static device array) against msplat_hip_radix_sort, through the C ABI: two calls in a
row on one context, the second smaller, bit for bit equal keys and values; sizes beyond
its maximum are refused on the host before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SORT_MAX = 1 << 20  # MSPLAT_NOALLOC_SORT_MAX


def _sort(ctx, fn, keys, vals, bits):
    from msplat._abi import DeviceArray, check
    dk, dv = DeviceArray.from_numpy(ctx, keys), DeviceArray.from_numpy(ctx, vals)
    check(fn(ctx.ptr, dk.ptr, dv.ptr, keys.size, bits))
    ctx.sync()
    return dk.numpy(), dv.numpy()


def test_noalloc_sort_equals_the_allocating_sort(ctx):
    from msplat._abi import lib
    L = lib()
    rng = np.random.default_rng(90)
    for n, bits in ((SORT_MAX, 45), (70_001, 30)):  # the second call smaller than the first
        keys = rng.integers(0, 2 ** 47, n, dtype=np.uint64)
        keys[: n // 3] &= np.uint64(0xFFF)  # many ties
        vals = np.arange(n, dtype=np.int32)
        want = _sort(ctx, L.msplat_hip_radix_sort, keys, vals, bits)
        got = _sort(ctx, L.msplat_hip_radix_sort_noalloc, keys, vals, bits)
        np.testing.assert_array_equal(got[0], want[0])
        np.testing.assert_array_equal(got[1], want[1])
        mask = np.uint64((1 << bits) - 1)
        order = np.argsort(keys & mask, kind="stable")
        np.testing.assert_array_equal(got[1], order.astype(np.int32))


def test_noalloc_sort_refuses_sizes_beyond_its_scratch(ctx):
    from msplat._abi import MsplatError, check, lib
    L = lib()
    with pytest.raises(MsplatError, match="at most"):
        check(L.msplat_hip_radix_sort_noalloc(ctx.ptr, None, None, SORT_MAX + 1, 32))
