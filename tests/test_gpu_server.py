"""The serving loop: results alone equal results in the mix; admission by budget.

Trees of every kind (octree LeafCount / NFCount, Morton / Hilbert, binary;
FP32 / FP64) are built and updated through one Server with all requests in
flight at once, then each tree's request sequence is replayed alone on a
plain Tree: every exported array must be identical (tolerance 0).
"""
import numpy as np
import pytest

from adaptive_trees_amd import (BINARY, HILBERT, LEAF_COUNT, MORTON, NF_COUNT, OCTREE, DeviceArray, Server, Tree,
                                TreeError, make_config)
from pairs import ARRAYS, jitter, plummer, uniform

pytestmark = pytest.mark.gpu

SPECS = [
    (OCTREE, np.float32, dict(sfc_kind=MORTON, max_particles_per_leaf=64), 50_000),
    (OCTREE, np.float64, dict(sfc_kind=HILBERT, split_criterion=NF_COUNT, max_particles_per_leaf=4096), 30_000),
    (BINARY, np.float32, dict(max_particles_per_leaf=32, box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)), 80_000),
    (BINARY, np.float64, dict(max_particles_per_leaf=16, views=(1, 1, 1)), 20_000),
    (OCTREE, np.float32, dict(sfc_kind=HILBERT, max_particles_per_leaf=16, views=(1, 1, 1)), 5_000),
    (OCTREE, np.float64, dict(sfc_kind=MORTON, max_particles_per_leaf=64), 100_000),
    (OCTREE, np.float32, dict(max_particles_per_leaf=8), 100),
    (BINARY, np.float32, dict(max_particles_per_leaf=4), 7),
]
UPDATES = 3


def snapshots(k, n, real):
    base = (plummer if k % 2 else uniform)(n, seed=300 + k)
    snaps = [base]
    for u in range(UPDATES):
        snaps.append(jitter(snaps[-1], 2e-3, seed=400 + 10 * k + u))
    return [[DeviceArray(np.ascontiguousarray(a, dtype=real)) for a in s] for s in snaps]


def copy_into(dst, src):
    from adaptive_trees_amd import _ffi
    for d, s in zip(dst, src):
        _ffi.hip_lib().call("memcpy", d.ptr, s.ptr, s.nbytes, 3)


@pytest.mark.parametrize("graphs", [True, False])
def test_results_alone_equal_results_in_the_mix(graphs):
    server = Server(hbm_budget_bytes=200 << 30, workers=4, max_batch=8, update_graphs=graphs)
    trees = []
    for k, (kind, real, cfg, n) in enumerate(SPECS):
        snaps = snapshots(k, n, real)
        coords = [DeviceArray(max(n, 1) * np.dtype(real).itemsize) for _ in range(3)]
        copy_into(coords, snaps[0])
        trees.append((kind, real, cfg, n, snaps, coords))
    ids, reqs = [], []
    for kind, real, cfg, n, snaps, coords in trees:
        tid, rid = server.submit_build(*coords, n, make_config(**cfg), real=real, kind=kind)
        ids.append(tid)
        reqs.append(rid)
    for u in range(1, UPDATES + 1):  # every update is queued before any result is read
        for tid, t in zip(ids, trees):
            reqs.append(server.submit_update(tid, new=t[4][u]))
    print("submitted", len(reqs), "requests", flush=True)
    for r in reqs:
        assert server.wait(r) == 0
        s, b, d = server.times(r)
        assert 0 <= s <= b <= d
    st = server.stats()
    assert st["completed"] == len(reqs)
    for tid, (kind, real, cfg, n, snaps, coords) in zip(ids, trees):
        alone_coords = [DeviceArray(max(n, 1) * np.dtype(real).itemsize) for _ in range(3)]
        copy_into(alone_coords, snaps[0])
        alone = Tree(*alone_coords, n, make_config(**cfg), real=real, kind=kind)
        alone.build()
        for u in range(1, UPDATES + 1):
            copy_into(alone_coords, snaps[u])
            alone.update()
            while alone.doBuffersNeedResize():
                alone.resizeBuffers()
                alone.update()
        mine = server.tree(tid)
        sa, sb = mine.scalars(), alone.scalars()
        for f in ("num_particles", "num_active_particles", "num_nodes", "num_leaves", "leaf_capacity"):
            assert getattr(sa, f) == getattr(sb, f), (tid, f)
        for name in ARRAYS:
            if name == "nearFieldCounts_d" and (kind == BINARY or sa.split_criterion == LEAF_COUNT):
                continue
            a, b = mine.get(name), alone.get(name)
            assert a.shape == b.shape and np.array_equal(a, b), (tid, name)
    server.close()


def test_admission_waits_for_budget():
    n = 100_000
    est = n * 25 + (n // 8 + 64) * 48 + (1 << 20)  # the server's per-tree estimate (no views)
    server = Server(hbm_budget_bytes=2 * est + 1000, workers=2)
    coords = [[DeviceArray(np.ascontiguousarray(a, dtype=np.float32)) for a in uniform(n, seed=500 + k)]
              for k in range(5)]
    builds = [server.submit_build(*c, n, make_config(max_particles_per_leaf=32), real=np.float32) for c in coords]
    assert server.stats()["waiting_builds"] >= 3
    done = []
    for tid, rid in builds:
        assert server.wait(rid) == 0
        done.append(tid)
        if len(done) >= 2:  # release the oldest to admit the next build
            assert server.wait(server.submit_release(done.pop(0))) == 0
    st = server.stats()
    assert st["bytes_peak"] <= 2 * est + 1000
    assert st["waiting_builds"] == 0
    server.close()


def test_request_errors():
    server = Server(hbm_budget_bytes=1 << 30, workers=1)
    with pytest.raises(TreeError):
        server.submit_update(12345)
    x = DeviceArray(np.zeros(10))
    tid, rid = server.submit_build(x, x, x, 10, make_config(max_particles_per_leaf=0))
    assert server.wait(rid) == 1  # AT_EINVAL from the tree constructor
    # a failed build keeps no tree and gives its bytes back
    assert server.stats()["bytes_in_use"] == 0
    assert server.wait(server.submit_update(tid)) == 2  # AT_ELOGIC: tree failed to build
    with pytest.raises(TreeError):
        server.tree(tid).scalars()
    assert server.wait(server.submit_release(tid)) == 0
    with pytest.raises(TreeError):
        big = DeviceArray(8)
        server.submit_build(big, big, big, 10 ** 9, make_config())  # exceeds the budget
    # once a release is queued the tree takes no further requests
    coords = [DeviceArray(np.ascontiguousarray(a, dtype=np.float32)) for a in uniform(1000, seed=7)]
    tid, rid = server.submit_build(*coords, 1000, make_config(max_particles_per_leaf=32), real=np.float32)
    release = server.submit_release(tid)
    with pytest.raises(TreeError):
        server.submit_update(tid)
    assert server.wait(rid) == 0 and server.wait(release) == 0
    assert server.stats()["resident_trees"] == 0 and server.stats()["bytes_in_use"] == 0
    server.close()


def test_destroy_completes_builds_waiting_for_budget():
    import threading
    import time
    n = 20_000
    est = n * 25 + (n // 8 + 64) * 48 + (1 << 20)
    server = Server(hbm_budget_bytes=est + 1000, workers=1)
    coords = [[DeviceArray(np.ascontiguousarray(a, dtype=np.float32)) for a in uniform(n, seed=600 + k)]
              for k in range(2)]
    cfg = make_config(max_particles_per_leaf=32)
    first, rid0 = server.submit_build(*coords[0], n, cfg, real=np.float32)
    second, rid1 = server.submit_build(*coords[1], n, cfg, real=np.float32)
    held = server.submit_update(second)  # held behind the build that cannot be admitted
    assert server.wait(rid0) == 0
    assert server.stats()["waiting_builds"] == 1
    seen = {}

    def waiter():
        seen["status"] = server.wait(rid1)

    t = threading.Thread(target=waiter)
    t.start()
    time.sleep(0.5)  # the waiter is blocked inside at_server_wait
    assert not server.poll(rid1) and not server.poll(held)
    server.close()
    t.join(timeout=30)
    assert not t.is_alive()
    assert seen["status"] == 2  # AT_ELOGIC: destroyed before admission


def pair_chains(n, seed):
    """Pairs of nearly coincident particles: with one particle per leaf, chains
    of empty leaves down to the deepest level (builds beyond a group's
    reservation for the device-driven passes)."""
    rng = np.random.default_rng(seed)
    base = rng.random((n // 2, 3))
    pts = np.concatenate([base, base + 1e-7 * rng.random((n // 2, 3))])[:n]
    return [np.ascontiguousarray(pts[:, a]) for a in range(3)]


# classes of several trees each: their builds and updates share launches
GROUPED = [
    (OCTREE, np.float32, dict(sfc_kind=MORTON, max_particles_per_leaf=64), [3000, 50_000, 120_000, 7, 4_500_000]),
    (OCTREE, np.float64, dict(sfc_kind=HILBERT, split_criterion=NF_COUNT, max_particles_per_leaf=4096),
     [20_000, 60_000, 9_000]),
    (OCTREE, np.float32, dict(sfc_kind=HILBERT, max_particles_per_leaf=16, views=(1, 1, 1)), [5_000, 40_000]),
    (OCTREE, np.float64, dict(sfc_kind=MORTON, max_particles_per_leaf=1), [400, 1000, 64]),
    (BINARY, np.float32, dict(max_particles_per_leaf=32), [30_000, 80_000]),
]


@pytest.mark.parametrize("workers", [1, 3])
def test_grouped_requests_equal_results_alone(workers):
    """Builds and updates of octrees of one configuration run as groups (one
    segmented sort and multi-tree passes over trees that keep their own
    buffers); every tree equals its requests run alone on a plain Tree,
    including bucket-1 pair chains that outgrow the group build's
    reservation, trees with views, a tree too large for a group and binary
    trees in the same rounds."""
    trees = []
    for c, (kind, real, cfg, sizes) in enumerate(GROUPED):
        for j, n in enumerate(sizes):
            seed = 1000 + 10 * c + j
            if cfg["max_particles_per_leaf"] == 1:
                base = pair_chains(n, seed)
                sigma = 1e-9
            else:
                base = (plummer if j % 2 else uniform)(n, seed=seed)
                sigma = 2e-3
            snaps = [base]
            for u in range(UPDATES):
                snaps.append(jitter(snaps[-1], sigma, seed=seed + 500 + u))
            dsnaps = [[DeviceArray(np.ascontiguousarray(a, dtype=real)) for a in s] for s in snaps]
            coords = [DeviceArray(max(n, 1) * np.dtype(real).itemsize) for _ in range(3)]
            copy_into(coords, dsnaps[0])
            trees.append((kind, real, cfg, n, dsnaps, coords))
    print("grouped mix: %d trees, workers %d" % (len(trees), workers), flush=True)

    def estimate(n, cfg):  # the server's per-tree estimate
        views = any(cfg.get("views", (0, 0, 0)))
        return n * 25 + (n // 8 + 64) * 48 + (n * 8 + (n // 8) * 26 if views else 0) + (1 << 20)

    # a blocker tree holds the budget while every request is submitted; its
    # release admits all builds at once (each followed by its held updates),
    # so the first round of a worker holds all of its builds
    nb = 8_000_000
    budget = estimate(nb, {}) + 1000
    assert sum(estimate(t[3], t[2]) for t in trees) <= budget
    blocker = [DeviceArray(np.ascontiguousarray(a, dtype=np.float32)) for a in uniform(nb, seed=999)]
    with Server(hbm_budget_bytes=budget, workers=workers, max_batch=64) as server:
        bid, brid = server.submit_build(*blocker, nb, make_config(max_particles_per_leaf=64), real=np.float32)
        assert server.wait(brid) == 0
        ids, reqs = [], []
        for kind, real, cfg, n, snaps, coords in trees:
            tid, rid = server.submit_build(*coords, n, make_config(**cfg), real=real, kind=kind)
            ids.append(tid)
            reqs.append(rid)
        for u in range(1, UPDATES + 1):
            for tid, t in zip(ids, trees):
                reqs.append(server.submit_update(tid, new=t[4][u]))
        assert server.stats()["waiting_builds"] == len(trees)
        assert server.wait(server.submit_release(bid)) == 0
        for r in reqs:
            assert server.wait(r) == 0
        st = server.stats()
        print("grouped builds %d, grouped updates %d" % (st["grouped_builds"], st["grouped_updates"]), flush=True)
        assert st["completed"] == len(reqs) + 2
        if workers == 1:
            # every octree of up to 4M particles is in a group of its class
            assert st["grouped_builds"] == 12 and st["grouped_updates"] == 12 * UPDATES
        else:
            assert st["grouped_builds"] > 0 and st["grouped_updates"] > 0
        for tid, (kind, real, cfg, n, snaps, coords) in zip(ids, trees):
            alone_coords = [DeviceArray(max(n, 1) * np.dtype(real).itemsize) for _ in range(3)]
            copy_into(alone_coords, snaps[0])
            alone = Tree(*alone_coords, n, make_config(**cfg), real=real, kind=kind)
            alone.build()
            for u in range(1, UPDATES + 1):
                copy_into(alone_coords, snaps[u])
                alone.update()
                while alone.doBuffersNeedResize():
                    alone.resizeBuffers()
                    alone.update()
            mine = server.tree(tid)
            sa, sb = mine.scalars(), alone.scalars()
            for f in ("num_particles", "num_active_particles", "num_nodes", "num_leaves", "leaf_capacity",
                      "active_buffer"):
                assert getattr(sa, f) == getattr(sb, f), (tid, n, f)
            for name in ARRAYS:
                if name == "nearFieldCounts_d" and (kind == BINARY or sa.split_criterion == LEAF_COUNT):
                    continue
                a, b = mine.get(name), alone.get(name)
                assert a.shape == b.shape and np.array_equal(a, b), (tid, n, name)
            alone.close()
