"""The reference's protocol tests on the HIP engine (tests/tree_test.cu).

Escaped slots and reactivation through one captured update graph
(:411-504), binary graph == direct through resize, split, merge and
reactivation (:960-1011), view flags and late view requests (:1013-1082),
empty / all-escaped inputs (:938-958), invalid configurations (:1084-1094),
and the pass-by-pass binary build used for buckets the direct construction
does not take. Every state is compared bit for bit with the oracle.
"""
import ctypes

import numpy as np
import pytest

from adaptive_trees_amd import (BINARY, HILBERT, LEAF_COUNT, MORTON, NF_COUNT, OCTREE, DeviceArray,
                                InvalidArgument, LogicError, Tree, _ffi, make_config)
from oracle.binding import OracleTree
from pairs import ARRAYS

FMAX = np.finfo(np.float64).max
pytestmark = pytest.mark.gpu


def lib():
    return _ffi.hip_lib()


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        lib().call("stream_create", ctypes.byref(p))
        self.p = p.value

    def sync(self):
        lib().call("stream_synchronize", self.p)


class Graph:
    def __init__(self, tree, stream):
        g = ctypes.c_void_p()
        lib().call("graph_capture_update", tree.handle, stream.p, ctypes.byref(g))
        self.g, self.stream = g.value, stream

    def replay(self):
        lib().call("graph_launch", self.g, self.stream.p)
        self.stream.sync()

    def close(self):
        if self.g:
            lib().call("graph_destroy", self.g)
            self.g = None


def same(a, b, what, kind):
    sa, sb = a.scalars(), b.scalars()
    for f in ("num_particles", "num_active_particles", "num_nodes", "num_leaves", "leaf_capacity"):
        assert getattr(sa, f) == getattr(sb, f), "%s: %s %d vs %d" % (what, f, getattr(sa, f), getattr(sb, f))
    for name in ARRAYS:
        if name == "nearFieldCounts_d" and (kind == BINARY or sa.split_criterion == LEAF_COUNT):
            continue
        x, y = a.get(name), b.get(name)
        assert x.shape == y.shape and np.array_equal(x, y), "%s: %s differs" % (what, name)
    assert a.maxAchievedDepth() == b.maxAchievedDepth(), what


class Trio:
    """A direct GPU tree, a graph-captured GPU tree and the oracle, same inputs."""

    def __init__(self, points, kind, real=np.float64, **cfg):
        self.kind, self.real = kind, np.dtype(real)
        self.stream = Stream()
        self.host = [np.ascontiguousarray(points[:, a], dtype=real) for a in range(3)]
        self.dev = [DeviceArray(a) for a in self.host]
        n = len(points)
        c = make_config(stream=self.stream.p, **cfg)
        self.direct = Tree(*self.dev, n, c, real=real, kind=kind)
        self.captured = Tree(*self.dev, n, make_config(stream=self.stream.p, **cfg), real=real, kind=kind)
        self.oracle = OracleTree(*self.host, n, make_config(**cfg), real=real, kind=kind)
        for t in (self.direct, self.captured, self.oracle):
            t.build()
        self.graph = Graph(self.captured, self.stream)

    def upload(self, points):
        for a in range(3):
            self.host[a][:] = points[:, a]
            self.dev[a].upload(self.host[a])

    def step(self):
        """One update() on all three; returns the (agreed) needsResize."""
        self.direct.update()
        self.graph.replay()
        self.oracle.update()
        r = {self.direct.doBuffersNeedResize(), self.captured.doBuffersNeedResize(),
             self.oracle.doBuffersNeedResize()}
        assert len(r) == 1, "needsResize differs"
        return r.pop()

    def resize(self):
        self.graph.close()
        for t in (self.direct, self.captured, self.oracle):
            t.resizeBuffers()
        self.graph = Graph(self.captured, self.stream)

    def compare(self, what):
        same(self.direct, self.oracle, what + " direct", self.kind)
        same(self.captured, self.oracle, what + " captured", self.kind)


def slot_points(mask):
    """tree_test.cu:248-262: slot s active at (0.25|0.75)^3 by its bits, else escaped."""
    out = np.empty((8, 3))
    for s in range(8):
        if mask >> s & 1:
            out[s] = [0.75 if s & 1 else 0.25, 0.75 if s & 2 else 0.25, 0.75 if s & 4 else 0.25]
        else:
            out[s] = FMAX
    return out


@pytest.mark.parametrize("sfc", [MORTON, HILBERT])
@pytest.mark.parametrize("crit,bucket", [(LEAF_COUNT, 1), (NF_COUNT, 27)])
def test_captured_update_escape_and_reactivation(sfc, crit, bucket):
    t = Trio(slot_points(0xFF), OCTREE, sfc_kind=sfc, split_criterion=crit, max_particles_per_leaf=bucket,
             views=(1, 1, 1))
    t.compare("build")
    addr = t.captured.buffer_info("particleKeys_d")[0], t.captured.buffer_info("perm_d")[0]
    for mask in (0xFF, 0xFF ^ 0x08, 0x91, 0x00, 0x04, 0xFF, 0x00, 0xFF):
        print("mask", hex(mask), flush=True)
        t.upload(slot_points(mask))
        while t.step():
            t.resize()
        t.compare("mask %#x" % mask)
        na = bin(mask).count("1")
        assert t.captured.numActiveParticles() == na
        P = t.captured.get("particleKeys_d")
        assert all(int(k) == 2 ** 64 - 1 for k in P[na:])
        assert (t.captured.buffer_info("particleKeys_d")[0], t.captured.buffer_info("perm_d")[0]) == addr
    t.graph.close()


def reference_points(box, count=65):
    """tree_test.cu:556-569 arithmetic point set."""
    i = np.arange(count)
    f = np.stack([((i * 173) % 1024) / 1024.0, ((i * 317) % 1024) / 1024.0, ((i * 563) % 1024) / 1024.0], 1)
    lo = np.array(box[0::2])
    hi = np.array(box[1::2])
    return lo + f * (hi - lo)


@pytest.mark.parametrize("iterations", [1, 2])
def test_binary_graph_matches_direct_through_resize_split_merge_reactivation(iterations):
    box = (0.0, 1.0, 0.0, 1.25, 0.0, 1.75)
    max_depth = 9
    empty = np.full((65, 3), FMAX)
    t = Trio(empty, BINARY, max_depth=max_depth, max_particles_per_leaf=1, update_iterations=iterations, box=box,
             views=(1, 1, 1))
    distributed = reference_points(box)
    partial = distributed.copy()
    partial[::3] = FMAX
    clustered = np.tile([0.25, 0.25, 0.25], (65, 1))
    resized = False
    for k, pts in enumerate((distributed, partial, clustered, empty, distributed)):
        print("point set", k, flush=True)
        t.upload(pts)
        for p in range(4 * max_depth + 16):
            grow = t.step()
            t.compare("set %d pass %d" % (k, p))
            if grow:
                resized = True
                t.resize()
    assert resized
    t.graph.close()


@pytest.mark.parametrize("views", [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1)])
@pytest.mark.parametrize("slots", [0, 9, 19])
def test_binary_view_flags(views, slots):
    box = (0.0, 1.0, -2.0, 0.0, 0.0, 8.0)
    pts = reference_points(box, slots) if slots == 19 else np.full((slots, 3), FMAX)
    if slots == 19:
        pts[5] = FMAX
    host = [np.ascontiguousarray(pts[:, a]) for a in range(3)]
    dev = [DeviceArray(a) for a in host]
    # tree_test.cu:542-548: the reference's test configuration uses bucket 3
    g = Tree(*dev, slots, make_config(max_particles_per_leaf=3, box=box, views=views), kind=BINARY)
    o = OracleTree(*host, slots, make_config(max_particles_per_leaf=3, box=box, views=views), kind=BINARY)
    for t in (g, o):
        t.build()
        t.computeViews()
        assert not t.doBuffersNeedResize()
    same(g, o, "views %s slots %d" % (views, slots), BINARY)


def test_binary_late_view_requests():
    box = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    pts = reference_points(box, 31)
    pts[7] = FMAX
    host = [np.ascontiguousarray(pts[:, a]) for a in range(3)]
    dev = [DeviceArray(a) for a in host]
    g = Tree(*dev, 31, make_config(max_particles_per_leaf=3), kind=BINARY)
    o = OracleTree(*host, 31, make_config(max_particles_per_leaf=3), kind=BINARY)
    for t in (g, o):
        t.build()
    for views in ((1, 0, 0), (1, 1, 0), (1, 1, 1)):
        for t in (g, o):
            t.computeViews(views)
            assert t.doBuffersNeedResize()
            t.resizeBuffers()
            t.computeViews(views)
            assert not t.doBuffersNeedResize()
        same(g, o, "late views %s" % (views,), BINARY)
    assert g.maxAchievedDepth() > 0
    esc = np.full(31, FMAX)
    for d in dev:
        d.upload(esc)
    for h in host:
        h[:] = esc
    for t in (g, o):
        t.update()
        attempts = 0
        while t.doBuffersNeedResize():
            attempts += 1
            assert attempts < 64
            t.resizeBuffers()
            t.update()
        assert len(t.get("leafCounts_d")) == 1
        assert t.maxAchievedDepth() == 0
        t.computeViews((1, 1, 1))
        assert not t.doBuffersNeedResize()
    same(g, o, "all escaped", BINARY)


@pytest.mark.parametrize("kind", [OCTREE, BINARY])
def test_empty_and_all_escaped(kind):
    for n in (0, 1, 7):
        pts = np.full((n, 3), FMAX)
        host = [np.ascontiguousarray(pts[:, a]) for a in range(3)]
        dev = [DeviceArray(a) for a in host]
        g = Tree(*dev, n, make_config(views=(1, 1, 1)), kind=kind)
        o = OracleTree(*host, n, make_config(views=(1, 1, 1)), kind=kind)
        for t in (g, o):
            t.build()
            t.update()
            t.computeViews()
            t.doBuffersNeedResize()
        same(g, o, "n=%d escaped" % n, kind)
        assert g.numActiveParticles() == 0


def test_invalid_configurations_raise():
    x = DeviceArray(np.zeros(16))
    bad = [
        (BINARY, dict(max_depth=64)),
        (BINARY, dict(box=(0.0, 0.0, 0.0, 1.0, 0.0, 1.0))),
        (BINARY, dict(box=(0.0, float("inf"), 0.0, 1.0, 0.0, 1.0))),
        (BINARY, dict(split_criterion=NF_COUNT)),
        (BINARY, dict(sfc_kind=HILBERT)),
        (BINARY, dict(update_iterations=0)),
        (OCTREE, dict(max_particles_per_leaf=0)),
        (OCTREE, dict(update_iterations=0)),
        (OCTREE, dict(max_depth=-2)),
    ]
    for kind, cfg in bad:
        with pytest.raises(InvalidArgument):
            Tree(x, x, x, 16, make_config(**cfg), kind=kind)
    t = Tree(x, x, x, 16, make_config(), kind=OCTREE)
    with pytest.raises(LogicError):
        t.update()
    with pytest.raises(LogicError):
        t.computeViews()
    b = Tree(x, x, x, 16, make_config(), kind=BINARY)
    b.build()
    with pytest.raises(LogicError):
        b.get("nearFieldCounts_d")


@pytest.mark.parametrize("bucket", [1500, 5000])
def test_binary_large_bucket_pass_build(bucket):
    """Buckets above the direct construction's limit build pass by pass."""
    rng = np.random.default_rng(90 + bucket)
    n = 400_000
    pts = np.concatenate([0.5 + 0.01 * rng.standard_normal((n // 2, 3)), rng.random((n - n // 2, 3))]).clip(0, 1)
    host = [np.ascontiguousarray(pts[:, a], dtype=np.float32) for a in range(3)]
    dev = [DeviceArray(a) for a in host]
    g = Tree(*dev, n, make_config(max_particles_per_leaf=bucket, views=(1, 1, 1)), real=np.float32, kind=BINARY)
    o = OracleTree(*host, n, make_config(max_particles_per_leaf=bucket, views=(1, 1, 1)), real=np.float32,
                   kind=BINARY)
    for t in (g, o):
        t.build()
        t.computeViews()
        t.doBuffersNeedResize()
    same(g, o, "bucket %d" % bucket, BINARY)


@pytest.mark.parametrize("bucket", [2, 7, 8, 9, 63, 64, 65, 300, 1024])
def test_binary_direct_build_window_edges(bucket):
    """The direct construction's window maxima: register windows for buckets
    up to 64 (with a separate form below 8), a sparse table above, up to the
    direct limit of 1024; a clustered set so that windows mix depths."""
    rng = np.random.default_rng(300 + bucket)
    n = 150_000 + bucket
    pts = np.concatenate([0.3 + 0.002 * rng.standard_normal((n // 3, 3)), rng.random((n - n // 3, 3))]).clip(0, 1)
    host = [np.ascontiguousarray(pts[:, a], dtype=np.float32) for a in range(3)]
    dev = [DeviceArray(a) for a in host]
    cfg = make_config(max_particles_per_leaf=bucket, box=(0.0, 2.0, 0.0, 1.0, 0.0, 1.0))
    g = Tree(*dev, n, cfg, real=np.float32, kind=BINARY)
    o = OracleTree(*host, n, cfg, real=np.float32, kind=BINARY)
    for t in (g, o):
        t.build()
        t.doBuffersNeedResize()
    same(g, o, "direct bucket %d" % bucket, BINARY)


@pytest.mark.parametrize("sfc", [MORTON, HILBERT])
@pytest.mark.parametrize("crit,bucket", [(LEAF_COUNT, 8), (NF_COUNT, 300)])
def test_octree_update_rejected_for_capacity(sfc, crit, bucket):
    """Octree updates whose passes overflow the leaf capacity
    (octree_rebalance.cuh:1177-1273: scan[cap] > cap rejects the whole pass
    and sets needsResize until resizeBuffers): the tree is built with every
    slot escaped (one leaf, a capacity of a few), then the particles arrive
    uniformly, so the split passes of the following updates outgrow the
    capacity again and again as it grows by a quarter per resize; then they
    contract into a small cluster and spread out again. Direct, captured and
    oracle trees agree after every call."""
    rng = np.random.default_rng(77 + bucket + sfc)
    base = rng.random((4000, 3))
    t = Trio(np.full((4000, 3), FMAX), OCTREE, sfc_kind=sfc, split_criterion=crit, max_particles_per_leaf=bucket)
    t.compare("build")
    resized = 0
    for k, f in enumerate((1.0, 0.2, 0.01, 0.003, 1.0)):
        print("contraction", f, flush=True)
        t.upload(0.3 + (base - 0.3) * f)
        for p in range(12):
            grow = t.step()
            t.compare("set %d pass %d" % (k, p))
            if grow:
                resized += 1
                t.resize()
    print("resizes", resized, flush=True)
    assert resized > 0
    t.graph.close()
