"""HIP engine vs CPU oracle, bit for bit (needs an MI355X)."""
import numpy as np
import pytest

from adaptive_trees_amd import BINARY, HILBERT, MORTON, NF_COUNT, OCTREE
from pairs import Pair, jitter, plummer, uniform

pytestmark = pytest.mark.gpu

ALL_VIEWS = (True, True, True)


@pytest.mark.parametrize("real", [np.float64, np.float32])
@pytest.mark.parametrize("sfc", [MORTON, HILBERT])
def test_octree_leafcount_build_update(real, sfc):
    x, y, z = uniform(65536, 1)
    p = Pair(x, y, z, OCTREE, real, max_particles_per_leaf=64, sfc_kind=sfc, views=ALL_VIEWS)
    p.build()
    p.compare("build")
    c = (x, y, z)
    for f in range(3):
        c = jitter(c, 1e-3, 10 + f)
        p.set_coords(*c)
        p.update()
        p.compare("update %d" % f)


@pytest.mark.parametrize("sfc", [MORTON, HILBERT])
def test_octree_nfcount_plummer(sfc):
    x, y, z = plummer(50000, 2)
    p = Pair(x, y, z, OCTREE, np.float32, max_particles_per_leaf=4096, split_criterion=NF_COUNT,
             sfc_kind=sfc, views=ALL_VIEWS)
    p.build()
    p.compare("build")
    c = (x, y, z)
    for f in range(3):
        c = jitter(c, 1e-3, 20 + f)
        p.set_coords(*c)
        p.update()
        p.compare("update %d" % f)


# 4:1:1 splits x twice before the x/y/z cycle (one-axis prefix); 4:2:1 splits
# x, x, y first (a prefix mixing axes)
@pytest.mark.parametrize("box", [(0, 4, 0, 1, 0, 1), (0, 4, 0, 2, 0, 1)], ids=["4-1-1", "4-2-1"])
@pytest.mark.parametrize("real", [np.float64, np.float32])
def test_binary_anisotropic(real, box):
    x, y, z = uniform(40000, 3, box)
    p = Pair(x, y, z, BINARY, real, max_particles_per_leaf=32, box=box, views=ALL_VIEWS)
    p.build()
    p.compare("build")
    c = (x, y, z)
    for f in range(3):
        c = jitter(c, 1e-3, 30 + f, box)
        p.set_coords(*c)
        p.update()
        p.compare("update %d" % f)


def test_reference_fixture(fixture_npz):
    d = fixture_npz
    p = Pair(d["x0000"], d["y0000"], d["z0000"], OCTREE, np.float64, max_particles_per_leaf=64 * 64 * 27,
             split_criterion=NF_COUNT, views=ALL_VIEWS)
    p.build()
    p.compare("build")
    p.set_coords(d["x1000"], d["y1000"], d["z1000"])
    p.update()
    p.compare("update")
    for name in ("particleKeys_d", "perm_d", "cornerstone_d", "sfcBoxIndex_d", "particleBeginIndex_d",
                 "particleCounts_d", "hasBoxSplit_d", "boxDepth_d", "parentIndex_d", "childIndex_d",
                 "sfcParticleIndex_d", "levelOffset_d"):
        a = p.gpu.get(name)
        assert np.array_equal(a, d[name].astype(a.dtype)), name
    assert p.gpu.maxAchievedDepth() == int(d["scalar_maxAchievedDepth"])
