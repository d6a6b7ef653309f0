"""Batched build/update of many octrees equals each tree built on its own.

A batch of uneven trees (0 to 150k particles, uniform or Plummer, some with
escaped slots) is built with shared launches, then updated through a
captured graph after the particles move; every member's P, Perm, K, N (and
NF) and state must equal a plain Tree of the same particles driven by the
same calls, and a few members are also checked against the oracle. Batches
of more than 2^20 particles in total sort with the MSD levels (every member
is a parent bucket of level 1); their members, each below 2^20, are
compared with plain trees that sort with LSD passes.
"""
import ctypes

import numpy as np
import pytest

from adaptive_trees_amd import (HILBERT, LEAF_COUNT, MORTON, NF_COUNT, OCTREE, Batch, DeviceArray, Tree, _ffi,
                                make_config)
from oracle.binding import OracleTree
from pairs import jitter, plummer, uniform

pytestmark = pytest.mark.gpu
SIZES = [0, 1, 7, 1000, 65, 20000, 150_000, 3, 50_000, 999, 4096, 2]
SIZES_MSD = [600_000, 0, 300_001, 5, 250_000, 70_000, 2047, 2049, 100_000]
NAMES = ("particleKeys_d", "perm_d", "cornerstone_d", "leafCounts_d")


def members(seed, real, sizes=SIZES):
    rng = np.random.default_rng(seed)
    esc = np.finfo(real).max  # the escaped coordinate of the tree's Real type
    out = []
    for i, n in enumerate(sizes):
        x, y, z = (uniform if i % 2 else plummer)(n, seed=seed + i) if n else (np.zeros(0),) * 3
        x, y, z = (np.array(a, dtype=real) for a in (x, y, z))
        if n > 10 and i % 3 == 0:  # some escaped slots
            idx = rng.choice(n, size=n // 10, replace=False)
            x[idx] = y[idx] = z[idx] = esc
        out.append([x, y, z])
    return out


def stream():
    p = ctypes.c_void_p()
    _ffi.hip_lib().call("stream_create", ctypes.byref(p))
    return p.value


def same(member, tree, what, nf):
    sa, sb = member.scalars(), tree.scalars()
    for f in ("num_particles", "num_active_particles", "num_leaves", "leaf_capacity"):
        assert getattr(sa, f) == getattr(sb, f), "%s: %s %d vs %d" % (what, f, getattr(sa, f), getattr(sb, f))
    for name in NAMES + (("nearFieldCounts_d",) if nf else ()):
        a, b = member.get(name), tree.get(name)
        assert a.shape == b.shape and np.array_equal(a, b), "%s: %s differs" % (what, name)


@pytest.mark.parametrize("real", [np.float32, np.float64])
@pytest.mark.parametrize("sfc,crit,bucket", [(MORTON, LEAF_COUNT, 64), (HILBERT, LEAF_COUNT, 16),
                                              (MORTON, NF_COUNT, 4096)])
def test_batch_equals_individual_trees(real, sfc, crit, bucket):
    run_batch(real, sfc, crit, bucket, SIZES)


@pytest.mark.parametrize("real,sfc,crit,bucket", [(np.float32, MORTON, LEAF_COUNT, 64),
                                                   (np.float64, HILBERT, LEAF_COUNT, 16),
                                                   (np.float32, HILBERT, NF_COUNT, 64)])
def test_msd_batch_equals_individual_trees(real, sfc, crit, bucket):
    assert sum(SIZES_MSD) >= 1 << 20
    run_batch(real, sfc, crit, bucket, SIZES_MSD)


def run_batch(real, sfc, crit, bucket, SIZES):
    s = stream()
    data = members(100 + bucket, real, SIZES)
    esc = np.finfo(real).max
    host = [[np.ascontiguousarray(a, dtype=real) for a in m] for m in data]
    dev = [[DeviceArray(a) for a in m] for m in host]
    cfg = dict(max_particles_per_leaf=bucket, sfc_kind=sfc, split_criterion=crit)
    batch = Batch([m[0] for m in dev], [m[1] for m in dev], [m[2] for m in dev], SIZES,
                  make_config(stream=s, **cfg), real=real)
    plain = [Tree(*m, n, make_config(**cfg), real=real, kind=OCTREE) for m, n in zip(dev, SIZES)]
    nf = crit == NF_COUNT
    print("batch build", real.__name__, sfc, crit, bucket, flush=True)
    batch.build()
    for t in plain:
        t.build()
    for k, t in enumerate(plain):
        same(batch.tree(k), t, "build member %d" % k, nf)
    graph = batch.capture_update(s)
    lib = _ffi.hip_lib()
    moved = [jitter([a.astype(np.float64) for a in m], 3e-3, seed=7 + k) if len(m[0]) else m
             for k, m in enumerate(data)]
    for rnd in range(3):
        for m, h, d, mv in zip(data, host, dev, moved):
            for a in range(3):
                v = np.where(m[a] == esc, esc, mv[a]) if rnd % 2 == 0 else m[a]
                h[a][:] = v
                d[a].upload(h[a])
        lib.call("graph_launch", graph, s)
        lib.call("stream_synchronize", s)
        for t in plain:
            t.update()
        flags = batch.needsResize()
        for k, t in enumerate(plain):
            assert bool(flags[k]) == t.doBuffersNeedResize(), "needsResize member %d" % k
        while flags.any():
            for k, t in enumerate(plain):
                if flags[k]:
                    t.resizeBuffers()
            batch.resizeBuffers()
            lib.call("graph_destroy", graph)
            graph = batch.capture_update(s)
            lib.call("graph_launch", graph, s)
            lib.call("stream_synchronize", s)
            for t in plain:
                t.update()
            flags = batch.needsResize()
            for k, t in enumerate(plain):
                assert bool(flags[k]) == t.doBuffersNeedResize()
        for k, t in enumerate(plain):
            same(batch.tree(k), t, "round %d member %d" % (rnd, k), nf)
    lib.call("graph_destroy", graph)
    # and the oracle, for two members, after a direct batch update
    for k in (3, 5):
        o = OracleTree(*[h.copy() for h in host[k]], SIZES[k], make_config(**cfg), real=real, kind=OCTREE)
        o.build()
        b = Batch([dev[k][0]], [dev[k][1]], [dev[k][2]], [SIZES[k]], make_config(**cfg), real=real)
        b.build()
        assert not b.needsResize().any()
        same(b.tree(0), o, "oracle member %d" % k, nf)


def test_batch_build_beyond_reservation():
    """Members whose build outgrows the leaves reserved for the device-driven
    passes (bucket 1, pairs of nearly coincident particles: chains of 7 empty
    leaves per level down to level 21) finish with host-driven passes; every
    member equals its plain tree and one the oracle."""
    rng = np.random.default_rng(31)
    sizes = [400, 0, 1000, 64, 600]
    host = []
    for n in sizes:
        base = rng.random((n // 2, 3))
        pts = np.concatenate([base, base + 1e-7 * rng.random((n // 2, 3))])[:n] if n else np.zeros((0, 3))
        host.append([np.ascontiguousarray(pts[:, a], dtype=np.float64) for a in range(3)])
    dev = [[DeviceArray(a) for a in m] for m in host]
    cfg = dict(max_particles_per_leaf=1, sfc_kind=MORTON, split_criterion=LEAF_COUNT)
    s = stream()
    batch = Batch([m[0] for m in dev], [m[1] for m in dev], [m[2] for m in dev], sizes,
                  make_config(stream=s, **cfg), real=np.float64)
    print("batch build beyond reservation", flush=True)
    batch.build()
    plain = [Tree(*m, n, make_config(**cfg), real=np.float64, kind=OCTREE) for m, n in zip(dev, sizes)]
    for t in plain:
        t.build()
    leaves = [t.scalars().num_leaves for t in plain]
    assert leaves[2] > 8 * sizes[2] + 4160, "the test needs a member beyond its reservation: %s" % leaves
    for k, t in enumerate(plain):
        same(batch.tree(k), t, "member %d" % k, False)
    o = OracleTree(*[h.copy() for h in host[2]], sizes[2], make_config(**cfg), real=np.float64, kind=OCTREE)
    o.build()
    same(batch.tree(2), o, "oracle member 2", False)
    # a second build of the same batch reuses the grown buffers
    batch.build()
    for k, t in enumerate(plain):
        same(batch.tree(k), t, "rebuild member %d" % k, False)
