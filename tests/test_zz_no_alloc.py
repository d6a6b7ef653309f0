"""at_sort_pairs_noalloc (the sort whose workspace is a fixed device array of
the library) against at_sort_pairs on the same inputs, bit for bit.

Two calls in a row on one stream, the second with a smaller input, so the
second call runs over what the first left in the shared array; both with
slot values (the permutation the trees store) and with carried values. Sizes
beyond the array's limit fail before any launch.
"""
import ctypes

import numpy as np
import pytest

from adaptive_trees_amd import DeviceArray, _ffi

MAX_KEYS = 1 << 20


def _keys(rng, n):
    k = rng.integers(0, 2**63, n, dtype=np.uint64)
    k[rng.random(n) < 0.3] >>= np.uint64(40)  # heavy ties in the low digits
    k[rng.random(n) < 0.02] = np.uint64(2**64 - 1)  # escaped-slot keys
    return k


def _sort(name, keys, vals, stream):
    lib = _ffi.hip_lib()
    dk = DeviceArray(keys)
    dv = DeviceArray(vals if vals is not None else np.full(len(keys), 7, np.uint32))
    lib.call(name, dk.ptr, dv.ptr, len(keys), 1 if vals is None else 0, stream)
    lib.call("stream_synchronize", stream)
    return dk.download(), dv.download()


@pytest.mark.gpu
@pytest.mark.parametrize("carried", [False, True], ids=["slots", "values"])
def test_noalloc_sort_equals_allocating_sort(carried):
    lib = _ffi.hip_lib()
    stream = ctypes.c_void_p()
    lib.call("stream_create", ctypes.byref(stream))
    try:
        rng = np.random.default_rng(2024 + carried)
        for n in (MAX_KEYS, 300_001, 999):
            print("noalloc sort n=%d carried=%d" % (n, carried), flush=True)
            keys = _keys(rng, n)
            vals = rng.integers(0, 2**32, n, dtype=np.uint32) if carried else None
            k_ref, v_ref = _sort("sort_pairs", keys, vals, stream)
            k_na, v_na = _sort("sort_pairs_noalloc", keys, vals, stream)
            assert np.array_equal(k_na, k_ref), "keys differ at n=%d" % n
            assert np.array_equal(v_na, v_ref), "values differ at n=%d" % n
            order = np.argsort(keys, kind="stable")
            assert np.array_equal(k_na, keys[order])
            assert np.array_equal(v_na, vals[order] if carried else order.astype(np.uint32))
    finally:
        lib.call("stream_destroy", stream)


@pytest.mark.gpu
def test_noalloc_sort_rejects_sizes_beyond_its_array():
    lib = _ffi.hip_lib()
    dk = DeviceArray(np.zeros(4, np.uint64))
    dv = DeviceArray(np.zeros(4, np.uint32))
    # the size check runs before the pointers are used: nothing is launched
    with pytest.raises(_ffi.LengthError):
        lib.call("sort_pairs_noalloc", dk.ptr, dv.ptr, MAX_KEYS + 1, 1, None)
    with pytest.raises(_ffi.InvalidArgument):
        lib.call("sort_pairs_noalloc", None, None, 10, 1, None)
