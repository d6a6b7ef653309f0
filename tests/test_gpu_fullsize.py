"""Parity at the benchmark sizes: the HIP engine against the oracle (and, for
batches and serving, against plain trees) on the full workloads.

The protocol is the reference's regression test (tests/tree_test.cu:155-209:
build, update, compare every array), driven at:
  - 65,536 FP64 uniform octree, Morton, LeafCount 64 (build + 3 updates);
  - 2^24 FP32 Plummer octree (a = 0.05), Hilbert, NFCount 4096: build + 10
    updates of reflected Gaussian steps (sigma 1e-3), every array after every
    call, views after the build and after the last update;
  - 2^25 FP32 binary tree, 4:1:1 box, two Gaussian clusters + 20% uniform,
    LeafCount 32: build, then updates replayed from one captured graph on the
    resident positions and on moved ones;
  - 512 octrees of 1e3..2e6 particles (log-uniform), FP32 uniform/Plummer,
    Morton, LeafCount 64: one batch build + 5 graph-replayed update rounds,
    all 512 members against plain trees driven by the same calls and every
    16th member against the oracle;
  - a served mix of trees up to 8M particles (octree/binary, FP32/FP64,
    Morton/Hilbert, LeafCount/NFCount) with every request in flight at once,
    each tree against the same requests replayed alone.
All comparisons are exact.
"""
import ctypes

import numpy as np
import pytest

from adaptive_trees_amd import (BINARY, HILBERT, LEAF_COUNT, MORTON, NF_COUNT, OCTREE, Batch, DeviceArray, Server,
                                Tree, _ffi, make_config, workloads)
from oracle.binding import OracleTree
from pairs import ARRAYS

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(1200)]
TREE_ARRAYS = ("particleKeys_d", "perm_d", "cornerstone_d", "leafCounts_d", "nearFieldCounts_d")
VIEW_ARRAYS = tuple(a for a in ARRAYS if a not in TREE_ARRAYS)


def log(*a):
    print(*a, flush=True)


def new_stream():
    p = ctypes.c_void_p()
    _ffi.hip_lib().call("stream_create", ctypes.byref(p))
    return p.value


def compare(gpu, cpu, what, names, kind, nf):
    sg, sc = gpu.scalars(), cpu.scalars()
    # (which A/B half is active is not compared: the one-pass binary build
    # leaves its tree in half 0, the reference's pass loop wherever the last
    # pass put it; the getters return the same contents)
    for f in ("num_particles", "num_active_particles", "num_leaves", "leaf_capacity"):
        assert getattr(sg, f) == getattr(sc, f), "%s %s: gpu %d oracle %d" % (what, f, getattr(sg, f),
                                                                           getattr(sc, f))
    for name in names:
        if name == "nearFieldCounts_d" and (kind == BINARY or not nf):
            continue
        a, b = gpu.get(name), cpu.get(name)
        assert a.shape == b.shape, "%s %s: shape gpu %s oracle %s" % (what, name, a.shape, b.shape)
        if not np.array_equal(a, b):
            idx = np.nonzero(a != b)[0]
            raise AssertionError("%s %s: %d mismatches, first at %s: gpu %s oracle %s" % (
                what, name, len(idx), idx[:4], a[idx[:4]], b[idx[:4]]))
    if hasattr(gpu, "device_counters"):
        assert gpu.device_counters() == cpu.device_counters(), what


def resize_loop(gpu, cpu):
    """doBuffersNeedResize on both, then the reference's resize protocol
    (resizeBuffers + update until no resize is needed); returns the resizes."""
    resizes = 0
    while True:
        rg, rc = gpu.doBuffersNeedResize(), cpu.doBuffersNeedResize()
        assert rg == rc, "needsResize: gpu %s oracle %s" % (rg, rc)
        if not rg:
            return resizes
        resizes += 1
        assert resizes < 64
        gpu.resizeBuffers()
        cpu.resizeBuffers()
        gpu.update()
        cpu.update()


def test_c1_octree_65k_fp64_uniform():
    n = 65536
    pts = workloads.uniform_cube(n, seed=1, dtype=np.float64)
    host = [a.copy() for a in pts]
    dev = [DeviceArray(a) for a in host]
    cfg = dict(max_particles_per_leaf=64, sfc_kind=MORTON, views=(1, 1, 1))
    gpu = Tree(*dev, n, make_config(**cfg), real=np.float64, kind=OCTREE)
    cpu = OracleTree(*host, n, make_config(**cfg), real=np.float64, kind=OCTREE)
    gpu.build()
    cpu.build()
    compare(gpu, cpu, "C1 build", ARRAYS, OCTREE, False)
    cur = pts
    for step in range(3):
        cur = workloads.reflect_step(cur, 1e-3, seed=100 + step)
        for h, d, a in zip(host, dev, cur):
            h[:] = a
            d.upload(h)
        gpu.update()
        cpu.update()
        resize_loop(gpu, cpu)
        compare(gpu, cpu, "C1 update %d" % step, ARRAYS, OCTREE, False)


def test_c2_octree_16M_plummer_hilbert_nfcount():
    n = 1 << 24
    log("C2: generating %d Plummer particles" % n)
    pts = workloads.plummer(n, seed=2, dtype=np.float32)
    host = [a.copy() for a in pts]
    dev = [DeviceArray(a) for a in host]
    cfg = dict(max_particles_per_leaf=4096, split_criterion=NF_COUNT, sfc_kind=HILBERT)
    gpu = Tree(*dev, n, make_config(**cfg), real=np.float32, kind=OCTREE)
    cpu = OracleTree(*host, n, make_config(**cfg), real=np.float32, kind=OCTREE)
    gpu.build()
    log("C2: gpu build done")
    cpu.build()
    log("C2: oracle build done, %d leaves" % cpu.scalars().num_leaves)
    compare(gpu, cpu, "C2 build", TREE_ARRAYS, OCTREE, True)
    gpu.computeViews((1, 1, 1))
    cpu.computeViews((1, 1, 1))
    compare(gpu, cpu, "C2 build views", VIEW_ARRAYS, OCTREE, True)
    cur, resizes = pts, 0
    for step in range(10):
        cur = workloads.reflect_step(cur, 1e-3, seed=200 + step)
        for h, d, a in zip(host, dev, cur):
            h[:] = a
            d.upload(h)
        gpu.update()
        cpu.update()
        resizes += resize_loop(gpu, cpu)
        compare(gpu, cpu, "C2 update %d" % step, TREE_ARRAYS, OCTREE, True)
        log("C2: update %d equal (%d leaves, %d resizes so far)" % (step, gpu.scalars().num_leaves, resizes))
    gpu.computeViews((1, 1, 1))
    cpu.computeViews((1, 1, 1))
    compare(gpu, cpu, "C2 last update views", VIEW_ARRAYS, OCTREE, True)
    assert gpu.maxAchievedDepth() == cpu.maxAchievedDepth()


def test_c3_binary_32M_two_clusters_graph_updates():
    n = 1 << 25
    box = workloads.ANISO_BOX
    log("C3: generating %d particles" % n)
    pts = workloads.two_clusters(n, seed=3, dtype=np.float32)
    moved = workloads.reflect_step(pts, 1e-3, seed=33, box=box)
    host = [a.copy() for a in pts]
    dev = [DeviceArray(a) for a in host]
    s = new_stream()
    lib = _ffi.hip_lib()
    cfg = dict(max_particles_per_leaf=32, box=box)
    gpu = Tree(*dev, n, make_config(stream=s, **cfg), real=np.float32, kind=BINARY)
    cpu = OracleTree(*host, n, make_config(**cfg), real=np.float32, kind=BINARY)
    gpu.build()
    cpu.build()
    log("C3: builds done, %d leaves" % cpu.scalars().num_leaves)
    compare(gpu, cpu, "C3 build", TREE_ARRAYS, BINARY, False)
    graph = ctypes.c_void_p()
    lib.call("graph_capture_update", gpu.handle, s, ctypes.byref(graph))
    try:
        for what, coords in (("static", pts), ("moved", moved), ("moved again", moved), ("back", pts)):
            for h, d, a in zip(host, dev, coords):
                h[:] = a
                d.upload(h)
            lib.call("graph_launch", graph.value, s)
            lib.call("stream_synchronize", s)
            cpu.update()
            rg, rc = gpu.doBuffersNeedResize(), cpu.doBuffersNeedResize()
            assert rg == rc
            while rg:  # resize, re-capture, replay (the graph holds the old buffers)
                gpu.resizeBuffers()
                cpu.resizeBuffers()
                lib.call("graph_destroy", graph.value)
                lib.call("graph_capture_update", gpu.handle, s, ctypes.byref(graph))
                lib.call("graph_launch", graph.value, s)
                lib.call("stream_synchronize", s)
                cpu.update()
                rg, rc = gpu.doBuffersNeedResize(), cpu.doBuffersNeedResize()
                assert rg == rc
            compare(gpu, cpu, "C3 graph update (%s)" % what, TREE_ARRAYS, BINARY, False)
            log("C3: graph update (%s) equal, %d leaves" % (what, gpu.scalars().num_leaves))
    finally:
        lib.call("graph_destroy", graph.value)
    gpu.computeViews((1, 1, 1))
    cpu.computeViews((1, 1, 1))
    compare(gpu, cpu, "C3 views", VIEW_ARRAYS, BINARY, False)


def test_c4_batch_512_trees():
    trees, rounds, seed = 512, 5, 4
    sizes = workloads.batch_sizes(trees, seed=seed)
    kinds = workloads.batch_kinds(trees, seed=seed)
    snaps = []  # per member: rounds + 1 snapshots (host FP32)
    for k, n in enumerate(sizes):
        gen = workloads.plummer if kinds[k] else workloads.uniform_cube
        c = gen(int(n), seed=seed * 100003 + k, dtype=np.float32)
        m = workloads.reflect_step(c, 1e-3, seed=seed * 100019 + k)
        snaps.append((c, m))
    total = int(sizes.sum())
    log("C4: %d trees, %d particles" % (trees, total))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    s = new_stream()
    lib = _ffi.hip_lib()
    cfg = dict(max_particles_per_leaf=64, sfc_kind=MORTON)
    # one concatenated coordinate array per axis for the batch, another for the plain trees
    bdev = [DeviceArray(total * 4, np.float32) for _ in range(3)]
    pdev = [DeviceArray(total * 4, np.float32) for _ in range(3)]

    def load(r):
        for a in range(3):
            v = np.concatenate([snaps[k][r % 2][a] for k in range(trees)])
            bdev[a].upload(v)
            pdev[a].upload(v)

    def ptrs(arrs, a):
        return [arrs[a].ptr + 4 * int(off[k]) for k in range(trees)]

    load(0)
    batch = Batch(ptrs(bdev, 0), ptrs(bdev, 1), ptrs(bdev, 2), sizes, make_config(stream=s, **cfg),
                  real=np.float32)
    plain = [Tree(*(ptrs(pdev, a)[k] for a in range(3)), int(sizes[k]), make_config(**cfg), real=np.float32,
                  kind=OCTREE) for k in range(trees)]
    sample = list(range(0, trees, 16))
    oracle = {}
    for k in sample:
        h = [snaps[k][0][a].copy() for a in range(3)]
        oracle[k] = (OracleTree(*h, int(sizes[k]), make_config(**cfg), real=np.float32, kind=OCTREE), h)
    batch.build()
    for t in plain:
        t.build()
    for k, (o, h) in oracle.items():
        o.build()

    def check(what):
        for k, t in enumerate(plain):
            m = batch.tree(k)
            sa, sb = m.scalars(), t.scalars()
            for f in ("num_particles", "num_active_particles", "num_leaves", "leaf_capacity"):
                assert getattr(sa, f) == getattr(sb, f), "%s member %d %s" % (what, k, f)
            for name in ("particleKeys_d", "perm_d", "cornerstone_d", "leafCounts_d"):
                assert np.array_equal(m.get(name), t.get(name)), "%s member %d %s" % (what, k, name)
        for k, (o, h) in oracle.items():
            compare(batch.tree(k), o, "%s member %d vs oracle" % (what, k),
                    ("particleKeys_d", "perm_d", "cornerstone_d", "leafCounts_d"), OCTREE, False)

    check("C4 build")
    log("C4: build equal (512 vs plain, %d vs oracle)" % len(oracle))
    graph = batch.capture_update(s)
    try:
        for r in range(1, rounds + 1):
            load(r)
            for k, (o, h) in oracle.items():
                for a in range(3):
                    h[a][:] = snaps[k][r % 2][a]
            lib.call("graph_launch", graph, s)
            lib.call("stream_synchronize", s)
            for t in plain:
                t.update()
            for k, (o, h) in oracle.items():
                o.update()
            flags = batch.needsResize()
            for k, t in enumerate(plain):
                assert bool(flags[k]) == t.doBuffersNeedResize(), "round %d member %d needsResize" % (r, k)
            for k, (o, h) in oracle.items():
                assert bool(flags[k]) == o.doBuffersNeedResize()
            while flags.any():
                for k, t in enumerate(plain):
                    if flags[k]:
                        t.resizeBuffers()
                for k, (o, h) in oracle.items():
                    if flags[k]:
                        o.resizeBuffers()
                batch.resizeBuffers()
                lib.call("graph_destroy", graph)
                graph = batch.capture_update(s)
                lib.call("graph_launch", graph, s)
                lib.call("stream_synchronize", s)
                for t in plain:
                    t.update()
                for k, (o, h) in oracle.items():
                    o.update()
                flags = batch.needsResize()
                for k, t in enumerate(plain):
                    assert bool(flags[k]) == t.doBuffersNeedResize()
                for k, (o, h) in oracle.items():
                    assert bool(flags[k]) == o.doBuffersNeedResize()
            if r in (1, rounds):
                check("C4 round %d" % r)
                log("C4: round %d equal" % r)
    finally:
        lib.call("graph_destroy", graph)


def test_c5_served_mix_up_to_8M():
    sizes = [int(v) for v in workloads.batch_sizes(9, seed=5, lo=10_000, hi=8_000_000)] + [8_000_000]
    specs = []
    for k, n in enumerate(sizes):
        kind = OCTREE if k % 10 < 7 else BINARY
        real = np.float32 if k % 2 else np.float64
        cfg = dict(max_particles_per_leaf=64)
        if kind == OCTREE:
            cfg["sfc_kind"] = HILBERT if (k // 2) % 2 else MORTON
            if k % 3 == 0:
                cfg.update(split_criterion=NF_COUNT, max_particles_per_leaf=4096)
        specs.append((kind, real, cfg, int(n)))
    updates = 3
    log("C5: %d trees, sizes %s" % (len(specs), sizes))
    server = Server(hbm_budget_bytes=200 << 30, workers=4, max_batch=16)
    trees, ids, reqs = [], [], []
    for k, (kind, real, cfg, n) in enumerate(specs):
        base = (workloads.plummer if k % 2 else workloads.uniform_cube)(n, seed=700 + k, dtype=np.float64)
        snaps = [base]
        for u in range(updates):
            snaps.append(workloads.reflect_step(snaps[-1], 1e-3, seed=800 + 10 * k + u))
        dsnaps = [[DeviceArray(np.ascontiguousarray(a, dtype=real)) for a in sn] for sn in snaps]
        coords = [DeviceArray(n * np.dtype(real).itemsize) for _ in range(3)]
        for d, sv in zip(coords, dsnaps[0]):
            _ffi.hip_lib().call("memcpy", d.ptr, sv.ptr, sv.nbytes, 3)
        trees.append((kind, real, cfg, n, dsnaps, coords))
    for kind, real, cfg, n, dsnaps, coords in trees:
        tid, rid = server.submit_build(*coords, n, make_config(**cfg), real=real, kind=kind)
        ids.append(tid)
        reqs.append(rid)
    for u in range(1, updates + 1):
        for tid, t in zip(ids, trees):
            reqs.append(server.submit_update(tid, new=t[4][u]))
    for r in reqs:
        assert server.wait(r) == 0
    log("C5: %d requests served" % len(reqs))
    for tid, (kind, real, cfg, n, dsnaps, coords) in zip(ids, trees):
        ac = [DeviceArray(n * np.dtype(real).itemsize) for _ in range(3)]
        for d, sv in zip(ac, dsnaps[0]):
            _ffi.hip_lib().call("memcpy", d.ptr, sv.ptr, sv.nbytes, 3)
        alone = Tree(*ac, n, make_config(**cfg), real=real, kind=kind)
        alone.build()
        alone.doBuffersNeedResize()
        for u in range(1, updates + 1):
            for d, sv in zip(ac, dsnaps[u]):
                _ffi.hip_lib().call("memcpy", d.ptr, sv.ptr, sv.nbytes, 3)
            alone.update()
            while alone.doBuffersNeedResize():
                alone.resizeBuffers()
                alone.update()
        mine = server.tree(tid)
        sa, sb = mine.scalars(), alone.scalars()
        for f in ("num_particles", "num_active_particles", "num_leaves", "leaf_capacity"):
            assert getattr(sa, f) == getattr(sb, f), (tid, n, f)
        nf = sa.split_criterion == NF_COUNT
        for name in TREE_ARRAYS:
            if name == "nearFieldCounts_d" and (kind == BINARY or not nf):
                continue
            assert np.array_equal(mine.get(name), alone.get(name)), (tid, n, name)
        alone.close()
        log("C5: tree %d (%d particles) equal alone and in the mix" % (tid, n))
    server.close()
