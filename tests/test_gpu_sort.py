"""The device radix sort on its own (at_sort_pairs) against numpy's stable sort.

Sorts take the MSD levels with the LDS chunk sort from sort_msd_min_keys
pairs on (4096 by default; 2^20 and 4096 are both tested), the LSD passes
below it and when sort_msd is off. Distributions cover what the level
planning and the chunk sort branch on: uniform 64-bit keys, 63-bit keys with
the invalid key ~0 mixed in, a handful of distinct values, long duplicate
runs, clustered high bits, keys whose high bits repeat in runs for the chunk
fix-up (short runs for the insertion sort, runs of 17..64 for the wave rank
sort, and longer ones for each fallback), and sizes that are not
multiples of a tile. Results must be identical (keys and permutation).
"""
import zlib

import numpy as np
import pytest

from adaptive_trees_amd import DeviceArray, _ffi, tune

MASK63 = np.uint64((1 << 63) - 1)


def chunk_groups(n, rng, per_group, low_bits):
    """Keys that form LDS chunks of ~1500 keys agreeing above bit 48 (the
    MSD levels stop at the second digit), whose highest varying bit is 47: a
    group id g with ~1500 / per_group values ends at bit 47, random bits fill
    [0, low_bits), and the bits between are zero. The chunk's passes cover
    bits 32..47, so keys agree above bit 32 in runs of ~per_group (one run
    per g), and above bit 24 in runs of ~per_group / 2^(32 - low_bits)."""
    c = rng.integers(0, max(n // 1500, 1), n, dtype=np.uint64) << np.uint64(48)
    groups = max(1500 // per_group, 2)
    g = rng.integers(0, groups, n, dtype=np.uint64) << np.uint64(48 - groups.bit_length())
    return c | g | rng.integers(0, 2**low_bits, n, dtype=np.uint64)


def keys_of(kind, n, rng):
    if kind == "uniform":
        return rng.integers(0, 2**64 - 1, n, dtype=np.uint64, endpoint=True)
    if kind == "invalid_mix":
        k = rng.integers(0, 2**63, n, dtype=np.uint64)
        k[rng.random(n) < 0.05] = np.uint64(2**64 - 1)
        k[rng.random(n) < 0.01] = MASK63  # a valid key equal to the invalid key's low 63 bits
        return k
    if kind == "few":
        return rng.choice(np.array([3, 2**40 + 1, 2**62, 2**63 - 1], np.uint64), n)
    if kind == "dup_runs":
        vals = rng.integers(0, 2**63, n // 6000 + 2, dtype=np.uint64)
        return rng.permutation(np.repeat(vals, 6000)[:n])
    if kind == "clustered":
        hi = (rng.normal(0, 1, n) * 3 + 100).astype(np.int64).clip(0, 255).astype(np.uint64) << np.uint64(56)
        return hi | rng.integers(0, 2**56, n, dtype=np.uint64)
    if kind == "short_runs":
        # runs of ~5 keys that agree above bit 32 (the chunk's passes cover
        # bits 32..47): the insertion fix-up
        return chunk_groups(n, rng, 5, 32)
    if kind == "wave_runs":
        # runs of 17..64 keys that agree above bit 32: one wave per run
        return chunk_groups(n, rng, 30, 32)
    if kind == "mid_runs":
        # runs of ~200 that agree above bit 32 but are short above bit 24 (the
        # chunk sort's first fallback)
        return chunk_groups(n, rng, 190, 32)
    if kind == "long_runs":
        # runs of ~200 that agree above bit 24: both fix-ups fail, plain LSD
        return chunk_groups(n, rng, 190, 24)
    if kind == "two_values":
        # chunks of two distinct keys (bit 46 apart) in runs of ~750: the
        # fallback passes take only the window holding the varying bit
        return chunk_groups(n, rng, 750, 0)
    if kind == "low_bits":
        # chunks whose keys differ only below bit 16: every fix-up fails and
        # the chunk takes plain LSD passes
        hi = rng.integers(0, 64, n, dtype=np.uint64) << np.uint64(40)
        return hi | rng.integers(0, 2**16, n, dtype=np.uint64)
    raise ValueError(kind)


def gpu_sort(keys, vals=None):
    lib = _ffi.hip_lib()
    n = len(keys)
    dk = DeviceArray(keys)
    dv = DeviceArray(vals if vals is not None else np.zeros(n, np.uint32))
    lib.call("sort_pairs", dk.ptr, dv.ptr, n, 1 if vals is None else 0, None)
    return dk.download(), dv.download()


@pytest.fixture(params=[1, 0, 2], ids=["msd_from_2e20", "lsd", "msd_from_4096"])
def msd(request):
    """1: MSD levels from 2^20 keys (LSD passes below); 0: LSD passes only;
    2: MSD levels from 4096 keys (small sorts through the partition levels
    and the chunk sort)."""
    saved = tune.get()
    tune.set(sort_msd=1 if request.param else 0, sort_msd_min_keys=4096 if request.param == 2 else 1 << 20)
    yield request.param
    tune.set(**saved)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["uniform", "invalid_mix", "few", "dup_runs", "clustered", "short_runs",
                                  "wave_runs", "mid_runs", "long_runs", "two_values", "low_bits"])
@pytest.mark.parametrize("n", [1000, 300_001, (1 << 20) + 777, 3_000_017])
def test_sort_pairs_matches_stable_argsort(msd, kind, n):
    print("sort", msd, kind, n, flush=True)
    rng = np.random.default_rng(zlib.crc32(("%s %d" % (kind, n)).encode()))
    keys = keys_of(kind, n, rng)
    k, v = gpu_sort(keys)
    order = np.argsort(keys, kind="stable").astype(np.uint32)
    assert np.array_equal(v, order), "%s n=%d: permutation differs at %s" % (kind, n, np.nonzero(v != order)[0][:4])
    assert np.array_equal(k, keys[order])


@pytest.mark.gpu
def test_sort_pairs_carries_values(msd):
    rng = np.random.default_rng(91)
    n = (1 << 20) + 5
    keys = keys_of("clustered", n, rng)
    vals = rng.integers(0, 2**32, n, dtype=np.uint32)
    k, v = gpu_sort(keys, vals)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])


def test_sort_pairs_rejects_null_pointers():
    lib = _ffi.hip_lib()
    with pytest.raises(_ffi.InvalidArgument):
        lib.call("sort_pairs", None, None, 10, 1, None)


@pytest.fixture
def wide_digits():
    saved = tune.get()
    tune.set(sort_digit_bits=11, sort_msd_min_keys=1 << 20)
    yield
    tune.set(**saved)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["uniform", "invalid_mix", "few", "dup_runs", "clustered", "low_bits"])
@pytest.mark.parametrize("n", [1, 1000, 2047, 2048, 2049, 300_001, (1 << 20) - 1])
def test_wide_digit_lsd_matches_stable_argsort(wide_digits, kind, n):
    """Six passes of 11-bit digits (sorts below 2^20 keys), slot values."""
    print("wide sort", kind, n, flush=True)
    rng = np.random.default_rng(zlib.crc32(("wide %s %d" % (kind, n)).encode()))
    keys = keys_of(kind, n, rng)
    k, v = gpu_sort(keys)
    order = np.argsort(keys, kind="stable").astype(np.uint32)
    assert np.array_equal(v, order), "%s n=%d: permutation differs at %s" % (kind, n, np.nonzero(v != order)[0][:4])
    assert np.array_equal(k, keys[order])


@pytest.mark.gpu
def test_wide_digit_lsd_carries_values(wide_digits):
    rng = np.random.default_rng(92)
    n = 777_777
    keys = keys_of("dup_runs", n, rng)
    vals = rng.integers(0, 2**32, n, dtype=np.uint32)
    k, v = gpu_sort(keys, vals)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])
