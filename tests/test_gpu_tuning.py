"""Every radix-sort tuning the autotuner may pick gives the same trees.

The tile shape (threads x keys per thread), the number of key ranges and
the histogram source (digit bytes or keys) change only the schedule of the
sort; P, Perm, K and N must stay bit-exact against the oracle.
"""
import numpy as np
import pytest

from adaptive_trees_amd import BINARY, HILBERT, MORTON, OCTREE, _ffi
from pairs import Pair, plummer

TUNINGS = [
    (512, 12, 2048, 1, 0), (512, 12, 2048, 0, 0), (256, 8, 64, 1, 0), (256, 16, 4096, 0, 0),
    (1024, 8, 512, 1, 0), (1024, 8, 65536, 0, 0), (512, 16, 1024, 1, 0),
    (512, 8, 2048, 1, 1), (256, 12, 64, 0, 1), (512, 12, 4096, 1, 1), (256, 8, 100, 0, 1),
    (512, 16, 2048, 1, 1), (256, 16, 2048, 0, 1),  # prefetch requested where it is not instantiated
]
KEYS = ("sort_threads", "sort_items", "sort_ranges", "sort_digit_array", "sort_prefetch")


@pytest.fixture
def tuning():
    lib = _ffi.hip_lib()
    saved = {}
    for k in KEYS:
        import ctypes
        v = ctypes.c_int()
        lib.call("tune_get", k.encode(), ctypes.byref(v))
        saved[k] = v.value

    def apply(values):
        for k, v in zip(KEYS, values):
            lib.call("tune_set", k.encode(), int(v))

    yield apply
    apply([saved[k] for k in KEYS])


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", TUNINGS, ids=lambda c: "%dx%d_r%d_d%d_p%d" % c)
def test_tunings_are_bit_exact(tuning, cfg):
    tuning(cfg)
    print("tuning", cfg, flush=True)
    x, y, z = plummer(300_000, seed=71)
    # duplicates: ties must keep ascending slots
    x[::97], y[::97], z[::97] = 0.25, 0.5, 0.75
    p = Pair(x, y, z, kind=OCTREE, real=np.float32, max_particles_per_leaf=64, sfc_kind=HILBERT)
    p.build()
    p.compare("octree build %s" % (cfg,))
    b = Pair(x, y, z, kind=BINARY, real=np.float64, max_particles_per_leaf=16, sfc_kind=MORTON,
             views=(1, 1, 1))
    b.build()
    b.gpu.computeViews()
    b.cpu.computeViews()
    b.update()
    b.compare("binary build+update+views %s" % (cfg,))


def test_invalid_tunings_are_rejected():
    lib = _ffi.hip_lib()
    for k, v in (("sort_threads", 300), ("sort_items", 7), ("sort_ranges", 1), ("sort_digit_array", 2),
                 ("sort_prefetch", 3), ("sort_digit_bits", 9), ("sort_msd_min_keys", 2047),
                 ("sort_msd_min_keys", -1), ("no_such_key", 1)):
        with pytest.raises(_ffi.InvalidArgument):
            lib.call("tune_set", k.encode(), v)


@pytest.mark.gpu
def test_oversized_tile_is_rejected_at_build(tuning):
    tuning((1024, 16, 2048, 1, 0))
    x, y, z = plummer(50_000, seed=73)
    p = Pair(x, y, z, kind=OCTREE, real=np.float32, max_particles_per_leaf=64)
    with pytest.raises(_ffi.InvalidArgument, match="at most 8192"):
        p.gpu.build()


@pytest.mark.gpu
def test_autotuner_picks_a_valid_setting(tuning):
    from adaptive_trees_amd import tune
    saved = tune.get()
    print("autotune", flush=True)
    r = tune.autotune(1 << 20)
    assert r["sort_threads"] in (256, 512, 1024)
    assert r["sort_items"] in (8, 12, 16)
    assert r["sort_ranges"] in (2048, 4096, 8192, 16384, 32768)
    assert r["sort_digit_array"] in (0, 1)
    assert r["sort_digit_bits"] in (8, 11) and r["digit_bits_keys"] == (1 << 20) - 1
    assert r["lsd_ms_8_bits"] > 0 and r["lsd_ms_11_bits"] > 0
    assert r["sort_ms"] > 0
    assert {k: r[k] for k in tune.KEYS} == tune.get()
    # the picked setting is bit-exact as well
    x, y, z = plummer(200_000, seed=72)
    p = Pair(x, y, z, kind=OCTREE, real=np.float32, max_particles_per_leaf=32)
    p.build()
    p.compare("autotuned")
    tune.set(**saved)


@pytest.mark.gpu
@pytest.mark.parametrize("bits", [8, 11])
def test_digit_width_trees_bit_exact(bits):
    """Trees whose sorts take the LSD passes (below 2^20 keys, MSD threshold
    raised to 2^20) are the oracle's at either digit width: build, update,
    views."""
    from adaptive_trees_amd import tune
    saved = tune.get()
    tune.set(sort_digit_bits=bits, sort_msd_min_keys=1 << 20)
    try:
        print("digit bits", bits, flush=True)
        x, y, z = plummer(250_000, seed=74)
        x[::89], y[::89], z[::89] = 0.125, 0.5, 0.875
        for kind, real, bucket, sfc in ((OCTREE, np.float32, 64, HILBERT), (OCTREE, np.float64, 16, MORTON),
                                        (BINARY, np.float32, 32, MORTON)):
            p = Pair(x, y, z, kind=kind, real=real, max_particles_per_leaf=bucket, sfc_kind=sfc, views=(1, 1, 1))
            p.build()
            p.compare("build bits=%d kind=%d" % (bits, kind))
            p.update()
            p.compare("update bits=%d kind=%d" % (bits, kind))
    finally:
        tune.set(**saved)


@pytest.mark.gpu
@pytest.mark.parametrize("min_keys", [2048, 1 << 14])
def test_msd_threshold_trees_bit_exact(min_keys):
    """Trees whose sorts take the MSD levels and the LDS chunk sort at small
    sizes (thresholds below the default 4096 and above it) are the oracle's:
    build, update, views."""
    from adaptive_trees_amd import tune
    saved = tune.get()
    tune.set(sort_msd=1, sort_msd_min_keys=min_keys)
    try:
        assert tune.get()["sort_msd_min_keys"] == min_keys
        print("msd threshold", min_keys, flush=True)
        for n, seed in ((250_000, 74), (3000, 75), (20_000, 76)):
            x, y, z = plummer(n, seed=seed)
            x[::89], y[::89], z[::89] = 0.125, 0.5, 0.875
            for kind, real, bucket, sfc in ((OCTREE, np.float32, 64, HILBERT), (OCTREE, np.float64, 16, MORTON),
                                            (BINARY, np.float32, 32, MORTON)):
                p = Pair(x, y, z, kind=kind, real=real, max_particles_per_leaf=bucket, sfc_kind=sfc,
                         views=(1, 1, 1))
                p.build()
                p.compare("build n=%d min=%d kind=%d" % (n, min_keys, kind))
                p.update()
                p.compare("update n=%d min=%d kind=%d" % (n, min_keys, kind))
    finally:
        tune.set(**saved)
