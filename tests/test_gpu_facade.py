"""The header-only C++ facade (include/adaptive_octree/tree.hpp) end to end.

tests/cpp/facade_driver drives Tree<Real, Octree>, KDTree3D<Real> (moved into
Tree<Real, Binary>) through build(), an in-place move of the particles,
update() with the resize loop and computeViews(), for FP64 and FP32, and
checks the exception types of the reference facade (tree.hpp:110-130).
Every dumped getter must equal the oracle's on the same inputs.
"""
import os
import subprocess

import numpy as np
import pytest

from adaptive_trees_amd import BINARY, HILBERT, NF_COUNT, OCTREE, make_config
from oracle.binding import OracleTree
from pairs import jitter, plummer

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "facade_driver")
DTYPES = {"particleKeys_d": np.uint64, "perm_d": np.uint32, "cornerstone_d": np.uint64,
          "leafCounts_d": np.uint32, "nearFieldCounts_d": np.uint32, "sfcBoxIndex_d": np.uint64,
          "hasBoxSplit_d": np.uint8, "boxDepth_d": np.uint8, "parentIndex_d": np.int32,
          "childIndex_d": np.int32, "levelOffset_d": np.int32, "particleBeginIndex_d": np.uint32,
          "particleCounts_d": np.uint32, "sfcParticleIndex_d": np.uint64}


def oracle_run(x0, x1, real, kind, **cfg):
    h = [np.ascontiguousarray(a, dtype=real) for a in x0]
    t = OracleTree(*h, len(h[0]), make_config(**cfg), real=real, kind=kind)
    t.build()
    for a, b in zip(h, x1):
        a[:] = np.asarray(b, dtype=real)
    t.update()
    while t.doBuffersNeedResize():
        t.resizeBuffers()
        t.update()
    t.computeViews()
    t.doBuffersNeedResize()
    return t


@pytest.mark.gpu
def test_cpp_facade_matches_oracle(tmp_path):
    assert os.path.exists(DRIVER), "tests/cpp/facade_driver is not built"
    n = 40000
    x0 = plummer(n, seed=81)
    x1 = jitter(x0, 2e-3, seed=82)
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(np.uint64(n).tobytes())
        for a in list(x0) + list(x1):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    print("facade driver, n =", n, flush=True)
    r = subprocess.run([DRIVER, str(tmp_path / "in.bin"), str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "exceptions 6/6" in r.stdout
    for tag, real in (("f64", np.float64), ("f32", np.float32)):
        for kind, name, cfg in ((OCTREE, "octree", dict(max_particles_per_leaf=4096, split_criterion=NF_COUNT,
                                                         sfc_kind=HILBERT, views=(1, 1, 1))),
                                (BINARY, "binary", dict(max_particles_per_leaf=16, views=(1, 1, 1)))):
            t = oracle_run(x0, x1, real, kind, **cfg)
            for g, dt in DTYPES.items():
                if g == "nearFieldCounts_d" and kind == BINARY:
                    continue
                got = np.fromfile(tmp_path / ("%s_%s_%s.bin" % (tag, name, g)), dtype=dt)
                want = t.get(g)
                assert got.shape == want.shape, (tag, name, g, got.shape, want.shape)
                assert np.array_equal(got, want), (tag, name, g)
            s = open(tmp_path / ("%s_%s_scalars.txt" % (tag, name))).read().split()
            assert [int(v) for v in s] == [t.numParticles(), t.numActiveParticles(), t.numNodes(),
                                           t.maxAchievedDepth()]
