"""The no-allocation radix sort (sd_radix_sort_pairs_u32_noalloc) against the
allocating entry point (sd_radix_sort_pairs_u32) and the oracle.

Two calls in a row, the second on a smaller input, so that the second call
runs on the workspace the first one left behind. Results must be bit-identical.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from stardust_amd import _lib as L

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu():
    L.require_gpu()


def sort_alloc(keys, vals, b0, b1):
    n = len(keys)
    k, v = L.DeviceBuffer.from_array(keys), L.DeviceBuffer.from_array(vals)
    ka, va = L.DeviceBuffer(4 * n), L.DeviceBuffer(4 * n)
    nb = C.c_size_t()
    L.call("sd_radix_sort_temp_bytes", n, C.byref(nb))
    tmp = L.DeviceBuffer(nb.value)
    L.call("sd_radix_sort_pairs_u32", k.ptr, v.ptr, ka.ptr, va.ptr, n, b0, b1, tmp.ptr, nb.value, None)
    L.synchronize()
    return k.download(np.uint32, n), v.download(np.uint32, n)


def sort_noalloc(keys, vals, b0, b1):
    n = len(keys)
    k, v = L.DeviceBuffer.from_array(keys), L.DeviceBuffer.from_array(vals)
    ka, va = L.DeviceBuffer(4 * n), L.DeviceBuffer(4 * n)
    L.call("sd_radix_sort_pairs_u32_noalloc", k.ptr, v.ptr, ka.ptr, va.ptr, n, b0, b1, None)
    L.synchronize()
    return k.download(np.uint32, n), v.download(np.uint32, n)


def test_noalloc_sort_matches_allocating_sort(oracle):
    rng = np.random.default_rng(7)
    # second call smaller than the first; bit ranges with an odd and an even pass count
    for n, b0, b1 in ((3_000_017, 0, 24), (65_536, 0, 17), (1_000, 4, 13)):
        say("noalloc sort", n, b0, b1)
        keys = rng.integers(0, 1 << b1, n, dtype=np.uint64).astype(np.uint32)
        keys[: n // 7] = keys[0]  # a long run of equal keys (stability)
        vals = rng.permutation(n).astype(np.uint32)
        gk, gv = sort_noalloc(keys, vals, b0, b1)
        ak, av = sort_alloc(keys, vals, b0, b1)
        ek, ev = oracle.sort_pairs_u32(keys, vals, b0, b1)
        assert np.array_equal(gk, ak) and np.array_equal(gv, av)
        assert np.array_equal(gk, ek) and np.array_equal(gv, ev)


def test_noalloc_sort_rejects_sizes_beyond_the_limit():
    lib = L.load()
    n = 8388608 + 1
    st = lib.sd_radix_sort_pairs_u32_noalloc(None, None, None, None, n, 0, 32, None)
    assert st == L.SD_ERR_ARG
    assert b"exceeds" in lib.sd_last_error()
