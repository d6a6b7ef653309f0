"""Large sorts: MSD levels + LDS chunk sort against LSD passes and the oracle.

Sorts from the MSD threshold on (4096 (key, slot) pairs by default)
partition by the high digits first and finish buckets of up to 2048 keys in
LDS. Key distributions here drive
the level planning through its cases: clustered keys (three levels), few
distinct keys and long duplicate runs (all eight levels, equal-key runs
larger than a chunk), escaped particles (the invalid key ~0), anisotropic
binary keys. Trees must equal the oracle's with the MSD path on and off.
"""
import numpy as np
import pytest

from adaptive_trees_amd import BINARY, HILBERT, MORTON, OCTREE, tune
from pairs import Pair, plummer, uniform

N = (1 << 20) + 12345  # above the MSD threshold, not a multiple of the tile


@pytest.fixture(params=[1, 0], ids=["msd", "lsd"])
def msd(request):
    saved = tune.get()
    tune.set(sort_msd=request.param)
    yield request.param
    tune.set(**saved)


@pytest.mark.gpu
def test_msd_plummer_octree_build_update(msd):
    print("msd", msd, "plummer octree", flush=True)
    x, y, z = plummer(N, seed=81)
    p = Pair(x, y, z, kind=OCTREE, real=np.float32, max_particles_per_leaf=64, sfc_kind=HILBERT)
    p.build()
    p.compare("msd=%d build" % msd)
    rng = np.random.default_rng(82)
    x2 = np.clip(x + rng.normal(0, 1e-3, N), 0, 1)
    y2 = np.clip(y + rng.normal(0, 1e-3, N), 0, 1)
    z2 = np.clip(z + rng.normal(0, 1e-3, N), 0, 1)
    p.set_coords(x2, y2, z2)
    p.update()
    p.compare("msd=%d update" % msd)


@pytest.mark.gpu
def test_msd_duplicates_and_escaped(msd):
    print("msd", msd, "duplicates", flush=True)
    rng = np.random.default_rng(83)
    pts = np.array([[0.1, 0.2, 0.3], [0.7, 0.7, 0.7], [0.25, 0.5, 0.75]])
    pick = rng.integers(0, 3, N)
    x, y, z = (pts[pick, k].copy() for k in range(3))
    # a uniform tail and escaped particles (outside the box) -> invalid keys
    tail = rng.random(N) < 0.05
    x[tail], y[tail], z[tail] = rng.random(tail.sum()), rng.random(tail.sum()), rng.random(tail.sum())
    out = rng.random(N) < 0.01
    x[out] = 2.0  # clamped to the box face
    esc = rng.random(N) < 0.01  # escaped slots (all coordinates at the largest value): invalid keys
    big = np.finfo(np.float64).max
    x[esc], y[esc], z[esc] = big, big, big
    p = Pair(x, y, z, kind=OCTREE, real=np.float64, max_particles_per_leaf=64, sfc_kind=MORTON,
             max_depth=10)
    p.build()
    p.compare("msd=%d duplicates" % msd)


@pytest.mark.gpu
def test_msd_binary_anisotropic(msd):
    print("msd", msd, "binary", flush=True)
    x, y, z = uniform(N, seed=84, box=(0, 4, 0, 1, 0, 1))
    x[: N // 3] = 1.5 + 0.01 * np.random.default_rng(85).standard_normal(N // 3)
    p = Pair(x, y, z, kind=BINARY, real=np.float32, max_particles_per_leaf=32, box=(0, 4, 0, 1, 0, 1),
             views=(1, 1, 1))
    p.build()
    p.gpu.computeViews()
    p.cpu.computeViews()
    p.compare("msd=%d binary" % msd)
