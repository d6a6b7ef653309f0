"""Benchmark of the adaptive-tree build/update path on one MI355X.

Headline workload (the 32M-particle binary tree the 60% roofline target is
quoted on): N = 2^25 FP32 particles in a 4:1:1 box, 40% + 40% isotropic
Gaussian clusters (sigma = 0.02 of the long side, centres at 1/4 and 3/4 of
x) + 20% uniform background, longest-side bisection, LeafCount 32. A second
snapshot B is the first (A) moved by reflected Gaussian steps of sigma = 1e-3
of each side.

One step = build() on the resident snapshot, a device copy of the other
snapshot into the tree's borrowed coordinates (the particles move), and one
hipGraph-replayed update() that rebalances for the moved particles; the
snapshots alternate between steps. value = 2 N / step time (every particle
once through build and once through update). Coordinates are generated on
the host and uploaded once; the timed region touches HBM only.

Output: one JSON line -- the metric, the roofline of the dominant kernel,
the build's graded and compulsory fractions of the HBM roofline, phase
times, parity flags against the oracle, the other configurations (65,536
FP64 octree with views, 2^24 FP32 Plummer NFCount octree), the 512-tree
batch, the serving loop, and the CPU baseline of the oracle.
"""
import argparse
import ctypes
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from adaptive_trees_amd import (BINARY, HILBERT, MORTON, NF_COUNT, OCTREE, Batch, DeviceArray, Server,  # noqa: E402
                                Tree, make_config)
from adaptive_trees_amd import _ffi, tune, workloads  # noqa: E402

METRIC = "Mparticles/s build+update, % HBM roofline; batched vs plain speedup; p99 latency"
PEAK_HBM_GBS = 8000.0
PAIR_BYTES = 24  # read + write of one (8-byte key, 4-byte value) pair
# The dominant kernel: the LDS chunk sort of the MSD radix sort (msdLocalSortKernel),
# one launch per sort over all N pairs; its profile label and its name in rocprofv3 output.
DOMINANT_LABEL = "msd_local"
DOMINANT_KERNEL = "msdLocalSortKernel"
COPY_BYTES = 1 << 30  # calibration copy: reads and writes this many bytes
MSD_CHUNK = 2048      # buckets above this size are partitioned again by the next MSD level


def _lib():
    return _ffi.hip_lib()


def progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        _lib().call("stream_create", ctypes.byref(p))
        self.p = p.value

    def sync(self):
        _lib().call("stream_synchronize", self.p)


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        _lib().call("event_create", ctypes.byref(p))
        self.p = p.value

    def record(self, stream):
        _lib().call("event_record", self.p, stream.p)

    def ms_since(self, start):
        v = ctypes.c_float()
        _lib().call("event_elapsed_ms", start.p, self.p, ctypes.byref(v))
        return v.value


def profile_query(label):
    ms, n = ctypes.c_double(), ctypes.c_ulonglong()
    status = _lib().fn("profile_query")(label.encode(), ctypes.byref(ms), ctypes.byref(n))
    return (ms.value, n.value) if status == 0 else (0.0, 0)


def copy_into(dst, src, stream):
    for d, s in zip(dst, src):
        _lib().call("memcpy_async", d.ptr, s.ptr, s.nbytes, 3, stream.p)


# ---- the roofline's algorithmic bytes (per build / update / views) ----------------------
def build_bytes(n, s, p, leaves, nf=False):
    """Key generation N(3s+12), a sort of p pair passes plus one key read for
    the digit histograms N(24p+8), one streaming read of P and the writes of
    K and N (8N + 12L); NFCount adds one merged read of P and the NF writes."""
    b = n * (3 * s + 12) + n * (24 * p + 8) + 8 * n + 12 * leaves
    return b + (8 * n + 4 * leaves if nf else 0)


def update_bytes(n, s, p, leaves, passes, nf=False):
    """Key generation and sort as in a build, one read of P, and 40 bytes per
    leaf for each pass that changed the tree (NFCount: 8N + 8 per leaf and pass)."""
    b = n * (3 * s + 12) + n * (24 * p + 8) + 8 * n + 40 * passes * leaves
    return b + (8 * n + 8 * passes * leaves if nf else 0)


def view_bytes(n, nodes, leaves):
    return 16 * n + 26 * nodes + 12 * leaves


def frac(nbytes, ms):
    return round(nbytes / (ms * 1e-3) / (PEAK_HBM_GBS * 1e9), 4) if ms > 0 else None


def msd_pair_passes(P):
    """Pair passes of the MSD sort over these (sorted) keys: level 1 moves
    every key, level l >= 2 the keys of the buckets of more than 2048 keys
    that level l-1 left (equal 8(l-1)-bit prefixes), and the LDS chunk sort
    moves every key once more."""
    P = np.asarray(P, dtype=np.uint64)
    n = len(P)
    if n == 0:
        return 0.0
    moved = n
    for level in range(2, 9):
        prefix = P >> np.uint64(64 - 8 * (level - 1))
        cuts = np.flatnonzero(prefix[1:] != prefix[:-1]) + 1
        sizes = np.diff(np.concatenate([[0], cuts, [n]]))
        big = int(sizes[sizes > MSD_CHUNK].sum())
        if big == 0:
            break
        moved += big
    return round((moved + n) / n, 4)


def host_cores():
    """(threads of this process's CPU share, physical cores per lscpu)."""
    share = len(os.sched_getaffinity(0))
    phys = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        cps = re.search(r"Core\(s\) per socket:\s*(\d+)", out)
        sk = re.search(r"Socket\(s\):\s*(\d+)", out)
        if cps and sk:
            phys = int(cps.group(1)) * int(sk.group(1))
    except (OSError, subprocess.SubprocessError):
        pass
    return share, phys


def get_tree(t, names=("particleKeys_d", "perm_d", "cornerstone_d", "leafCounts_d")):
    return {k: t.get(k) for k in names}


def same_arrays(a, b):
    return all(np.array_equal(a[k], b[k]) for k in a)


# ---- headline: 32M binary tree ---------------------------------------------------------------
def headline(args):
    n = args.n
    box = workloads.ANISO_BOX
    tuning = tune.get() if args.no_autotune else tune.autotune(n)
    progress("sort tuning: %s" % json.dumps(tuning))
    A = workloads.two_clusters(n, seed=3, dtype=np.float32)
    B = workloads.reflect_step(A, 1e-3, seed=33, box=box)
    snaps = [[DeviceArray(a) for a in A], [DeviceArray(a) for a in B]]
    coords = [DeviceArray(a) for a in A]
    stream = Stream()
    cfg = make_config(max_particles_per_leaf=32, box=box, stream=stream.p)
    tree = Tree(*coords, n, cfg, real=np.float32, kind=BINARY)
    tree.build()
    graph = ctypes.c_void_p()
    _lib().call("graph_capture_update", tree.handle, stream.p, ctypes.byref(graph))
    addr = tree.buffer_info("cornerstone_d")[0]
    state = {"at": 0}  # which snapshot the coordinates hold

    def step():
        tree.build()
        state["at"] ^= 1
        copy_into(coords, snaps[state["at"]], stream)
        _lib().call("graph_launch", graph.value, stream.p)

    for _ in range(args.warmup):
        step()
    stream.sync()
    assert tree.buffer_info("cornerstone_d")[0] == addr, "build moved the captured buffers"
    resizes = 0
    while tree.doBuffersNeedResize():  # (never at this size: capacity after build covers the moves)
        resizes += 1
        tree.resizeBuffers()
        _lib().call("graph_destroy", graph.value)
        _lib().call("graph_capture_update", tree.handle, stream.p, ctypes.byref(graph))

    # the timed steps
    _lib().call("profile_reset")
    _lib().call("profile_enable", 1)
    e0, e1 = Event(), Event()
    stream.sync()
    e0.record(stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    e1.record(stream)
    stream.sync()
    wall = time.perf_counter() - t0
    _lib().call("profile_enable", 0)
    ms_per_step = wall * 1e3 / args.steps
    event_ms = e1.ms_since(e0) / args.steps
    while tree.doBuffersNeedResize():
        resizes += 1
        tree.resizeBuffers()

    def per_launch(label):
        ms, cnt = profile_query(label)
        return ms / max(cnt, 1), cnt

    dom_ms, dom_n = per_launch(DOMINANT_LABEL)
    key_ms, _ = per_launch("keygen")
    pass_ms, _ = per_launch("sort_pass")
    level_ms, level_n = per_launch("msd_level")
    dir_ms, dir_n = per_launch("bin_direct")
    achieved = PAIR_BYTES * n / (dom_ms * 1e-3) / 1e9 if dom_n else 0.0
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            t = json.load(f)
            if t.get("kernel") == DOMINANT_KERNEL:
                traffic = t.get("bytes_per_launch")

    # phases, event-timed one at a time outside the timed region (medians of 5)
    ea, eb = Event(), Event()

    def timed(fn):
        stream.sync()
        ea.record(stream)
        fn()
        eb.record(stream)
        stream.sync()
        return eb.ms_since(ea)

    def to(k):
        copy_into(coords, snaps[k], stream)
        state["at"] = k

    build_ms, moved_ms, static_ms, copy_ms, passes = [], [], [], [], []
    for _ in range(5):
        to(0)
        build_ms.append(timed(tree.build))
        copy_ms.append(timed(lambda: copy_into(coords, snaps[1], stream)))
        state["at"] = 1
        moved_ms.append(timed(lambda: _lib().call("graph_launch", graph.value, stream.p)))
        passes.append(tree.updatePasses())
        static_ms.append(timed(lambda: _lib().call("graph_launch", graph.value, stream.p)))
    build_med, moved_med = float(np.median(build_ms)), float(np.median(moved_ms))
    to(0)
    stream.sync()
    t0 = time.perf_counter()
    tree.build()
    stream.sync()
    build_wall_ms = (time.perf_counter() - t0) * 1e3

    # results of the parity protocol: build(A), then the graph update on B
    to(0)
    tree.build()
    stream.sync()
    after_build = get_tree(tree)
    leaves = len(after_build["leafCounts_d"])
    p_build = msd_pair_passes(after_build["particleKeys_d"])
    to(1)
    _lib().call("graph_launch", graph.value, stream.p)
    stream.sync()
    while tree.doBuffersNeedResize():
        resizes += 1
        tree.resizeBuffers()
        tree.update()
    after_update = get_tree(tree)
    moved_passes = int(np.median(passes))
    p_update = msd_pair_passes(after_update["particleKeys_d"])
    leaves_update = len(after_update["leafCounts_d"])

    # views of the same tree (reported separately; not part of the step)
    view_ms = []
    for _ in range(5):
        view_ms.append(timed(lambda: tree.computeViews((1, 1, 1))))
        while tree.doBuffersNeedResize():
            tree.resizeBuffers()
    view_med = float(np.median(view_ms[1:]))
    nodes = int(tree.scalars().num_nodes)

    # stream-copy ceiling of this GPU (event-timed copy kernel, 1 GiB read + 1 GiB written)
    src, dst = DeviceArray(COPY_BYTES), DeviceArray(COPY_BYTES)
    cps = []
    for _ in range(4):
        cps.append(timed(lambda: _lib().call("bench_copy", dst.ptr, src.ptr, COPY_BYTES, stream.p)))
    copy_ceiling = 2 * COPY_BYTES / (float(np.median(cps[1:])) * 1e-3) / 1e9
    del src, dst

    b_graded = build_bytes(n, 4, p_build, leaves)
    b_comp = build_bytes(n, 4, 1, leaves)
    u_graded = update_bytes(n, 4, p_update, leaves_update, moved_passes)
    out = {
        "metric": METRIC,
        "value": round(2 * n / (ms_per_step * 1e-3) / 1e6, 2),
        "unit": "Mparticles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": "binary longest-side tree, N=2^25 FP32, 4:1:1 box, two Gaussian clusters (sigma 0.02 of "
                        "the long side) + 20% uniform, LeafCount 32; step = build() + device copy of the moved "
                        "snapshot (reflected steps, sigma 1e-3 of each side) + graph-replayed update(), snapshots "
                        "alternating",
            "particles": n,
            "leaves": leaves,
            "tree": "binary",
            "bucket": 32,
            "sort_tuning": tuning,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "LDS chunk sort of the MSD radix sort (%s, one launch per sort over all N pairs)"
                      % DOMINANT_KERNEL,
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": traffic,
            "bytes_per_launch": PAIR_BYTES * n,
            "avg_launch_ms": round(dom_ms, 5),
        },
        "phases_ms": {
            "step_event": round(event_ms, 4),
            "build": round(build_med, 4),
            "build_wall": round(build_wall_ms, 4),
            "snapshot_copy": round(float(np.median(copy_ms)), 4),
            "update_moved_graph": round(moved_med, 4),
            "update_static_graph": round(float(np.median(static_ms)), 4),
            "keygen": round(key_ms, 4),
            "sort_level1_pass": round(pass_ms, 4),
            "sort_levels_2_to_8": round(level_ms * level_n / max(dom_n, 1), 4),
            "sort_lds_chunks": round(dom_ms, 4),
            "binary_construction": round(2 * dir_ms, 4),  # two launches (count, emit) per build
        },
        "rates_Mparticles_per_s": {
            "build": round(n / (build_med * 1e-3) / 1e6, 1),
            "update_moved": round(n / (moved_med * 1e-3) / 1e6, 1),
            "step_2N": round(2 * n / (ms_per_step * 1e-3) / 1e6, 1),
        },
        "build_roofline": {
            "sort_pair_passes_p": p_build,
            "bytes_graded": int(b_graded),
            "frac_graded": frac(b_graded, build_med),
            "bytes_compulsory_p1": int(b_comp),
            "frac_compulsory": frac(b_comp, build_med),
            "copy_ceiling_GBs": round(copy_ceiling, 1),
        },
        "update_roofline": {
            "sort_pair_passes_p": p_update,
            "passes_changed_I": moved_passes,
            "leaves": leaves_update,
            "bytes_graded": int(u_graded),
            "frac_graded": frac(u_graded, moved_med),
            "frac_compulsory": frac(update_bytes(n, 4, 1, leaves_update, moved_passes), moved_med),
        },
        "views": {
            "ms": round(view_med, 4),
            "nodes": nodes,
            "bytes": int(view_bytes(n, nodes, leaves_update)),
            "frac": frac(view_bytes(n, nodes, leaves_update), view_med),
        },
        "resizes": resizes,
    }
    _lib().call("graph_destroy", graph.value)
    tree.close()
    return out, (A, B, box, after_build, after_update)


# ---- other single-tree configurations ------------------------------------------------------------
def other_configs(check):
    """65,536 FP64 uniform octree (Morton, LeafCount 64): build median and
    views, checked against the oracle; 2^24 FP32 Plummer octree (a = 0.05,
    Hilbert, NFCount 4096): build median of 10 after warm-up, then 10
    updates of reflected Gaussian steps (sigma 1e-3), the build checked
    against the oracle."""
    stream = Stream()
    ea, eb = Event(), Event()

    def timed(fn):
        stream.sync()
        ea.record(stream)
        fn()
        eb.record(stream)
        stream.sync()
        return eb.ms_since(ea)

    out = {}
    n = 65536
    x, y, z = workloads.uniform_cube(n, seed=1, dtype=np.float64)
    dev = [DeviceArray(a) for a in (x, y, z)]
    t = Tree(*dev, n, make_config(max_particles_per_leaf=64, stream=stream.p), real=np.float64, kind=OCTREE)
    times = [timed(t.build) for _ in range(25)]
    views = []
    for _ in range(10):
        views.append(timed(lambda: t.computeViews((1, 1, 1))))
        while t.doBuffersNeedResize():
            t.resizeBuffers()
    leaves, nodes = int(t.scalars().num_leaves), int(t.scalars().num_nodes)
    exact = None
    if check:
        from oracle.binding import OracleTree
        o = OracleTree(x.copy(), y.copy(), z.copy(), n, make_config(max_particles_per_leaf=64, views=(1, 1, 1)),
                       real=np.float64, kind=OCTREE)
        o.build()
        exact = all(np.array_equal(t.get(k), o.get(k)) for k in
                    ("particleKeys_d", "perm_d", "cornerstone_d", "leafCounts_d", "sfcBoxIndex_d", "parentIndex_d",
                     "childIndex_d", "particleCounts_d", "sfcParticleIndex_d"))
    out["octree_65k_fp64_uniform"] = {
        "build_ms_median": round(float(np.median(times[5:])), 4), "leaves": leaves, "nodes": nodes,
        "views_ms_median": round(float(np.median(views[2:])), 4),
        "views_frac": frac(view_bytes(n, nodes, leaves), float(np.median(views[2:]))),
        "bit_exact_vs_oracle": exact, "note": "launch-bound (about 16 MB of traffic): reported, not graded"}
    t.close()
    progress("65k octree done")

    n = 1 << 24
    pts = workloads.plummer(n, seed=2, dtype=np.float32)
    dev = [DeviceArray(a) for a in pts]
    cfg = dict(max_particles_per_leaf=4096, split_criterion=NF_COUNT, sfc_kind=HILBERT)
    t = Tree(*dev, n, make_config(stream=stream.p, **cfg), real=np.float32, kind=OCTREE)
    builds = [timed(t.build) for _ in range(13)]
    build_med = float(np.median(builds[3:]))
    leaves_b = int(t.scalars().num_leaves)
    p_b = msd_pair_passes(t.get("particleKeys_d"))
    exact = None
    if check:
        from oracle.binding import OracleTree
        o = OracleTree(*[a.copy() for a in pts], n, make_config(**cfg), real=np.float32, kind=OCTREE)
        o.build()
        exact = all(np.array_equal(t.get(k), o.get(k)) for k in
                    ("particleKeys_d", "perm_d", "cornerstone_d", "leafCounts_d", "nearFieldCounts_d"))
        del o
    upd, passes, resizes, cur = [], [], 0, pts
    for step in range(10):
        cur = workloads.reflect_step(cur, 1e-3, seed=200 + step)
        for d, a in zip(dev, cur):
            d.upload(a)
        upd.append(timed(t.update))
        passes.append(t.updatePasses())
        while t.doBuffersNeedResize():
            resizes += 1
            t.resizeBuffers()
            t.update()
    leaves_u = int(t.scalars().num_leaves)
    p_u = msd_pair_passes(t.get("particleKeys_d"))
    upd_med = float(np.median(upd))
    I = int(np.median(passes))
    out["octree_16M_fp32_plummer_hilbert_nf4096"] = {
        "build_ms_median_of_10": round(build_med, 3), "build_ms_first": round(builds[0], 3),
        "update_ms_median": round(upd_med, 3), "update_ms_max": round(float(np.max(upd)), 3),
        "resizes": resizes, "leaves_after_build": leaves_b, "leaves": leaves_u,
        "sort_pair_passes_p": p_b, "update_passes_changed_I": I,
        "build_frac_graded": frac(build_bytes(n, 4, p_b, leaves_b, nf=True), build_med),
        "build_frac_compulsory": frac(build_bytes(n, 4, 1, leaves_b, nf=True), build_med),
        "update_frac_graded": frac(update_bytes(n, 4, p_u, leaves_u, I, nf=True), upd_med),
        "update_frac_compulsory": frac(update_bytes(n, 4, 1, leaves_u, I, nf=True), upd_med),
        "Mparticles_per_s_build": round(n / (build_med * 1e-3) / 1e6, 1),
        "Mparticles_per_s_update": round(n / (upd_med * 1e-3) / 1e6, 1),
        "bit_exact_build_vs_oracle": exact}
    t.close()
    progress("16M NF octree done")
    return out


# ---- batches -----------------------------------------------------------------------------------------
def batch_workload(trees, seed):
    """512 octree inputs: sizes log-uniform in [1e3, 2e6], uniform or Plummer per tree by a seed
    bit, FP32, plus one reflected-Gaussian moved snapshot (sigma = 1e-3 of the unit box)."""
    sizes = workloads.batch_sizes(trees, seed=seed)
    kinds = workloads.batch_kinds(trees, seed=seed)
    first, moved = [[], [], []], [[], [], []]
    for k, n in enumerate(sizes):
        gen = workloads.plummer if kinds[k] else workloads.uniform_cube
        c = gen(int(n), seed=seed * 100003 + k, dtype=np.float32)
        m = workloads.reflect_step(c, 1e-3, seed=seed * 100019 + k)
        for a in range(3):
            first[a].append(c[a])
            moved[a].append(m[a])
    return sizes, [np.concatenate(v) for v in first], [np.concatenate(v) for v in moved]


def batch_benchmark(trees=512, rounds=5, seed=4):
    """Batched path (one Batch: segmented build, then `rounds` graph-replayed updates) against
    the plain path (one Tree per input on one stream: build() + rounds x (update() +
    doBuffersNeedResize())) on the same device-resident particles. Each round first copies the
    next snapshot into the borrowed coordinate arrays (one device copy per axis for all trees,
    the same in both paths). Every member is compared with its plain tree afterwards."""
    lib = _lib()
    sizes, first, moved = batch_workload(trees, seed)
    progress("batched workload generated: %d trees, %d particles" % (trees, int(sizes.sum())))
    total = int(sizes.sum())
    off = np.concatenate([[0], np.cumsum(sizes)])
    cur_b = [DeviceArray(a) for a in first]
    cur_p = [DeviceArray(a) for a in first]
    snaps = [[DeviceArray(a) for a in first], [DeviceArray(a) for a in moved]]
    stream = Stream()
    cfg = dict(max_particles_per_leaf=64)

    def ptrs(cur, a):
        return [cur[a].ptr + 4 * int(off[k]) for k in range(trees)]

    def load(cur, r):
        copy_into(cur, snaps[r % 2], stream)

    resizes = {"plain": 0, "batch": 0}
    plain = [Tree(*(ptrs(cur_p, a)[k] for a in range(3)), int(sizes[k]), make_config(stream=stream.p, **cfg),
                  real=np.float32, kind=OCTREE) for k in range(trees)]

    def run_plain():
        load(cur_p, 0)
        for t in plain:
            t.build()
        for r in range(1, rounds + 1):
            load(cur_p, r)
            for t in plain:
                t.update()
            for t in plain:
                while t.doBuffersNeedResize():
                    resizes["plain"] += 1
                    t.resizeBuffers()
                    t.update()
        stream.sync()

    batch = Batch(ptrs(cur_b, 0), ptrs(cur_b, 1), ptrs(cur_b, 2), sizes, make_config(stream=stream.p, **cfg),
                  real=np.float32)
    graph = [None]

    def run_batch():
        load(cur_b, 0)
        batch.build()
        graph[0] = graph[0] or batch.capture_update(stream.p)
        for r in range(1, rounds + 1):
            load(cur_b, r)
            lib.call("graph_launch", graph[0], stream.p)
            while batch.needsResize().any():
                resizes["batch"] += 1
                batch.resizeBuffers()
                lib.call("graph_destroy", graph[0])
                graph[0] = batch.capture_update(stream.p)
                lib.call("graph_launch", graph[0], stream.p)
        stream.sync()

    run_plain()  # warm-up: allocations reach their steady state
    run_batch()
    progress("batched/plain warm-up done")
    resizes = {"plain": 0, "batch": 0}
    t0 = time.perf_counter()
    run_plain()
    plain_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    run_batch()
    batch_s = time.perf_counter() - t0
    exact = all(np.array_equal(batch.tree(k).get(n), plain[k].get(n))
                for k in range(trees) for n in ("particleKeys_d", "perm_d", "cornerstone_d", "leafCounts_d"))

    def build_s(fn, reps=5):
        times = []
        for _ in range(reps):
            load(cur_b, 0)
            load(cur_p, 0)
            stream.sync()
            t0 = time.perf_counter()
            fn()
            stream.sync()
            times.append(time.perf_counter() - t0)
        return float(np.median(times))

    batch_build = build_s(batch.build)
    plain_build = build_s(lambda: [t.build() for t in plain])
    work = total * (rounds + 1)
    return {
        "workload": "%d octrees, sizes log-uniform [1e3, 2e6] (total %d), FP32 uniform/Plummer, Morton, "
                    "LeafCount 64; build + %d graph-replayed updates (sigma 1e-3 reflected steps)"
                    % (trees, total, rounds),
        "particles": total,
        "plain_s": round(plain_s, 4),
        "batched_s": round(batch_s, 4),
        "speedup": round(plain_s / batch_s, 3),
        "build_only_ms": {"batched": round(1e3 * batch_build, 2), "plain": round(1e3 * plain_build, 2),
                          "speedup": round(plain_build / batch_build, 2), "statistic": "median of 5"},
        "batched_Mparticles_per_s": round(work / batch_s / 1e6, 1),
        "plain_Mparticles_per_s": round(work / plain_s / 1e6, 1),
        "resizes": resizes,
        "bit_exact_all_members_vs_plain": bool(exact),
        "oracle_parity": "tests/test_gpu_fullsize.py::test_c4_batch_512_trees (every 16th member)",
    }


def batch_profile(trees=512, rounds=5, seed=4, repeats=3):
    """The batched path alone: `repeats` times a build() and `rounds` graph-replayed updates of the
    512-tree workload, each timed on its own (stream synchronised around it); medians over the
    repeats after the first."""
    lib = _lib()
    sizes, first, moved = batch_workload(trees, seed)
    off = np.concatenate([[0], np.cumsum(sizes)])
    cur = [DeviceArray(a) for a in first]
    snaps = [[DeviceArray(a) for a in first], [DeviceArray(a) for a in moved]]
    stream = Stream()
    batch = Batch(*([cur[a].ptr + 4 * int(off[k]) for k in range(trees)] for a in range(3)), sizes,
                  make_config(stream=stream.p, max_particles_per_leaf=64), real=np.float32)
    graph, builds, updates = None, [], []
    for rep in range(repeats):
        copy_into(cur, snaps[0], stream)
        stream.sync()
        t0 = time.perf_counter()
        batch.build()
        stream.sync()
        builds.append(time.perf_counter() - t0)
        graph = graph or batch.capture_update(stream.p)
        for r in range(1, rounds + 1):
            copy_into(cur, snaps[r % 2], stream)
            stream.sync()
            t0 = time.perf_counter()
            lib.call("graph_launch", graph, stream.p)
            stream.sync()
            updates.append(time.perf_counter() - t0)
            if batch.needsResize().any():
                raise RuntimeError("batch_profile: a member needs a resize")
    lib.call("graph_destroy", graph)
    return {"particles": int(sizes.sum()), "build_ms": round(1e3 * float(np.median(builds[1:])), 3),
            "update_ms": round(1e3 * float(np.median(updates[rounds:])), 3), "repeats": repeats, "rounds": rounds}


# ---- serving -------------------------------------------------------------------------------------------
def serving_trace(rate, requests, pool, seed):
    """Poisson arrivals at `rate`/s. The first 8 requests build resident trees; afterwards 30%
    build a new tree (the oldest resident tree is released beyond 16) and 70% update a random
    resident tree with its next snapshot. Tree inputs come from `pool` (sizes log-uniform in
    [1e4, 8e6]); each build draws the type (70% octree / 30% binary), the precision (50/50) and,
    for octrees, Hilbert/Morton and LeafCount 64 / NFCount 4096 (50/50 each)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=requests))
    trace, resident, next_id = [], [], 0
    for i in range(requests):
        if i < 8 or rng.random() < 0.3 or not resident:
            kind = OCTREE if rng.random() < 0.7 else BINARY
            real = np.float32 if rng.random() < 0.5 else np.float64
            cfg = dict(max_particles_per_leaf=64)
            if kind == OCTREE:
                cfg["sfc_kind"] = HILBERT if rng.random() < 0.5 else MORTON
                if rng.random() < 0.5:
                    cfg.update(split_criterion=NF_COUNT, max_particles_per_leaf=4096)
            trace.append(("build", t[i], next_id, int(rng.integers(len(pool))), kind, real, cfg))
            resident.append(next_id)
            next_id += 1
            if len(resident) > 16:
                trace.append(("release", t[i], resident.pop(0)))
        else:
            trace.append(("update", t[i], resident[int(rng.integers(len(resident)))]))
    return trace


def serving_benchmark(rates, requests, seed=5, workers=4, saturated=800):
    """The serving loop (worker streams, 200 GB budget) against serial plain calls on the same
    trace; latency = result ready - arrival, per request. Besides the Poisson rates, a saturated
    trace submits all `saturated` requests at once (capacity: requests/s of the whole trace); each
    of its passes runs twice and the second is reported."""
    sizes = workloads.batch_sizes(20, seed=seed, lo=10_000, hi=8_000_000)
    pool = []  # per entry: two snapshots, host float64
    for k, n in enumerate(sizes):
        a = (workloads.plummer if k % 2 else workloads.uniform_cube)(int(n), seed=700 + k, dtype=np.float64)
        pool.append((a, workloads.reflect_step(a, 1e-3, seed=800 + k)))
    progress("serving pool generated: %d inputs, %d particles" % (len(pool), int(sizes.sum())))
    dev_pool = {}

    def snapshot(p, real, which):
        key = (p, np.dtype(real).name, which)
        if key not in dev_pool:
            dev_pool[key] = [DeviceArray(np.ascontiguousarray(v, dtype=real)) for v in pool[p][which]]
        return dev_pool[key]

    def served_pass(trace, prep, info, rate):
        # closed before anything else on every exit path: its workers read the
        # device arrays of this function until their queues are drained
        with Server(hbm_budget_bytes=200 << 30, workers=workers, max_batch=16) as server:
            ids, turn, reqs = {}, {}, []
            start = time.perf_counter()
            for r, bufs in zip(trace, prep):
                delay = r[1] - (time.perf_counter() - start)
                if delay > 0:
                    time.sleep(delay)
                if r[0] == "build":
                    _, _, tid, p, kind, real, cfg = r
                    ids[tid], rid = server.submit_build(*bufs[0], len(pool[p][0][0]), make_config(**cfg),
                                                        real=real, kind=kind)
                    turn[tid] = 0
                    reqs.append((rid, len(pool[p][0][0])))
                elif r[0] == "update":
                    tid = r[2]
                    _, _, _, p, kind, real, cfg = info[tid]
                    turn[tid] += 1
                    reqs.append((server.submit_update(ids[tid], new=snapshot(p, real, turn[tid] % 2)),
                                 len(pool[p][0][0])))
                else:
                    reqs.append((server.submit_release(ids[r[2]]), 0))
            lat, first, last, work, failed = [], None, 0.0, 0, []
            for rid, n in reqs:
                st = server.wait(rid)
                if st != 0:
                    failed.append((rid, st))
                    continue
                s_ms, b_ms, d_ms = server.times(rid)
                lat.append(d_ms - s_ms)
                first = s_ms if first is None else min(first, s_ms)
                last = max(last, d_ms)
                work += n
            stats = server.stats()
        if failed:
            raise RuntimeError("serving rate %s: %d requests failed (first: id %d, status %d)"
                               % (rate, len(failed), failed[0][0], failed[0][1]))
        return lat, (last - first) / 1e3, work, stats

    def serial_pass(trace, prep, info):
        plain, turn2, slat = {}, {}, []
        start = time.perf_counter()
        for r, bufs in zip(trace, prep):
            arrival = r[1]
            now = time.perf_counter() - start
            if now < arrival:
                time.sleep(arrival - now)
            if r[0] == "build":
                _, _, tid, p, kind, real, cfg = r
                t = Tree(*bufs[1], len(pool[p][0][0]), make_config(**cfg), real=real, kind=kind)
                t.build()
                t.doBuffersNeedResize()
                plain[tid], turn2[tid] = (t, bufs[1]), 0
            elif r[0] == "update":
                tid = r[2]
                t, coords = plain[tid]
                _, _, _, p, kind, real, cfg = info[tid]
                turn2[tid] += 1
                for d, s in zip(coords, snapshot(p, real, turn2[tid] % 2)):
                    _lib().call("memcpy", d.ptr, s.ptr, s.nbytes, 3)
                t.update()
                while t.doBuffersNeedResize():
                    t.resizeBuffers()
                    t.update()
            else:
                plain.pop(r[2])[0].close()
            slat.append((time.perf_counter() - start - arrival) * 1e3)
        serial_s = time.perf_counter() - start - trace[0][1]
        for t, _ in plain.values():
            t.close()
        return slat, serial_s

    results = []
    # an untimed warm-up trace first: kernel modules load and the memory pools fill on first use
    runs = [("warm-up", 150)] + [(rate, requests) for rate in rates] + (
        [("saturated", saturated)] if saturated else [])
    for rate, count in runs:
        trace = serving_trace(1e9 if isinstance(rate, str) else rate, count, pool,
                              seed + (2000 if rate == "saturated" else 3000 if rate == "warm-up" else rate))
        prep = []
        for r in trace:
            if r[0] == "build":
                _, _, tid, p, kind, real, cfg = r
                n = len(pool[p][0][0])
                bufs = [[DeviceArray(n * np.dtype(real).itemsize) for _ in range(3)] for _ in range(2)]
                for b in bufs:
                    for d, s in zip(b, snapshot(p, real, 0)):
                        _lib().call("memcpy", d.ptr, s.ptr, s.nbytes, 3)
                prep.append(bufs)
                for w in (0, 1):
                    snapshot(p, real, w)
            else:
                prep.append(None)
        info = {r[2]: r for r in trace if r[0] == "build"}
        for _ in range(2 if rate == "saturated" else 1):
            lat, served_s, work, stats = served_pass(trace, prep, info, rate)
            slat, serial_s = serial_pass(trace, prep, info)
        if rate == "warm-up":
            continue
        n_req = len(trace)
        results.append({
            "rate_per_s": rate, "requests": n_req,
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "serial_p50_ms": round(float(np.percentile(slat, 50)), 3),
            "serial_p99_ms": round(float(np.percentile(slat, 99)), 3),
            "requests_per_s": round(n_req / served_s, 1),
            "Mparticles_per_s": round(work / served_s / 1e6, 1),
            "serial_requests_per_s": round(n_req / serial_s, 1),
            "served_over_serial": round(serial_s / served_s, 3),
            "batched_updates": int(stats["batched_updates"]), "grouped_builds": int(stats["grouped_builds"]),
            "grouped_updates": int(stats["grouped_updates"]), "peak_GB": round(stats["bytes_peak"] / 1e9, 3),
        })
        progress("serving rate %s done" % rate)
    return {"workload": "Poisson trace, %d requests per rate, sizes log-uniform [1e4, 8e6], 70%% octree / 30%% "
                        "binary, FP32/FP64, Hilbert/Morton, LeafCount 64 / NFCount 4096; saturated: %d requests "
                        "submitted at once (second of two passes); after an untimed warm-up trace of 150 requests; "
                        "%d worker streams, 200 GB budget, rounds of "
                        "up to 16 requests, octrees of one configuration grouped, no graph captures; serial = "
                        "plain Tree calls in arrival order on one thread"
                        % (requests, saturated, workers),
            "rates": results}


def serving_in_child(rates, requests, workers, saturated, hw_queues):
    """The serving leg in a child process (this script with --only-serve) whose HIP runtime opens
    `hw_queues` hardware queues (GPU_MAX_HW_QUEUES; HIP's default is 4 per process). With 8, the
    server's worker streams and the library's own streams do not share queues (saturated trace:
    1.86x serial requests/s against 1.47x with 4 queues). Progress lines pass through."""
    env = dict(os.environ, GPU_MAX_HW_QUEUES=str(hw_queues))
    cmd = [sys.executable, os.path.abspath(__file__), "--only-serve", "--serve-rates", ",".join(str(r) for r in rates),
           "--serve-requests", str(requests), "--serve-workers", str(workers), "--serve-saturated", str(saturated)]
    r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, text=True, timeout=1500)
    if r.returncode != 0:
        raise RuntimeError("serving leg failed (exit status %d)" % r.returncode)
    out = json.loads(r.stdout.strip().splitlines()[-1])["serving"]
    out["hw_queues"] = int(env["GPU_MAX_HW_QUEUES"])
    return out


# ---- CPU baseline ----------------------------------------------------------------------------------------
def cpu_baseline(A, B, box, gpu_build, gpu_update):
    """The oracle (oracle/, OpenMP C++) on the headline workload in host memory: build() on
    snapshot A then update() on B, 1 warm-up + 5 timed repeats, build and update timed
    separately (medians); the last repeat's trees are compared with the GPU's. It runs in a
    child process (this script with --cpu-child) so that its thread placement
    (OMP_PLACES=cores, OMP_PROC_BIND=close, which binds the starting thread too) leaves the
    GPU measurements' host threads alone; inputs and results pass through files."""
    import tempfile
    share, phys = host_cores()
    with tempfile.TemporaryDirectory(dir="/dev/shm" if os.path.isdir("/dev/shm") else None) as d:
        np.savez(os.path.join(d, "in.npz"), ax=A[0], ay=A[1], az=A[2], bx=B[0], by=B[1], bz=B[2],
                 box=np.array(box, dtype=np.float64))
        env = dict(os.environ, OMP_PLACES="cores", OMP_PROC_BIND="close")
        env.setdefault("OMP_NUM_THREADS", str(share))
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", d], env=env,
                           capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            raise RuntimeError("CPU baseline failed:\n" + r.stderr[-2000:])
        res = json.loads(r.stdout.strip().splitlines()[-1])
        got = np.load(os.path.join(d, "out.npz"))
        exact_build = all(np.array_equal(gpu_build[k], got["build_" + k]) for k in gpu_build)
        exact_update = all(np.array_equal(gpu_update[k], got["update_" + k]) for k in gpu_update)
    n = len(A[0])
    b, u = res["build_s"], res["update_s"]
    return {
        "value": round(2 * n / (b + u) / 1e6, 3), "unit": "Mparticles/s", "cores": res["threads"], "kind": "port",
        "physical_cores_lscpu": phys, "cpu_share_threads": share,
        "omp": {"OMP_NUM_THREADS": env.get("OMP_NUM_THREADS"), "OMP_PLACES": "cores", "OMP_PROC_BIND": "close"},
        "build_ms_median": round(b * 1e3, 2), "update_ms_median": round(u * 1e3, 2),
        "build_Mparticles_per_s": round(n / b / 1e6, 2), "update_Mparticles_per_s": round(n / u / 1e6, 2),
        "sample": "oracle build() on snapshot A + update() on the moved snapshot B of the full %d-particle "
                  "workload in host memory, 1 warm-up + 5 timed repeats (%.1f s of CPU work)"
                  % (n, res["total_s"]),
    }, exact_build, exact_update


def cpu_child(d):
    """The CPU baseline's child process (see cpu_baseline)."""
    from oracle.binding import OracleTree, num_threads
    z = np.load(os.path.join(d, "in.npz"))
    A, B = (z["ax"], z["ay"], z["az"]), (z["bx"], z["by"], z["bz"])
    box = tuple(float(v) for v in z["box"])
    n = len(A[0])
    h = [a.copy() for a in A]
    t = OracleTree(*h, n, make_config(max_particles_per_leaf=32, box=box), real=np.float32, kind=BINARY)
    tb, tu = [], []
    for rep in range(6):
        for dst, a in zip(h, A):
            dst[:] = a
        t0 = time.perf_counter()
        t.build()
        t1 = time.perf_counter()
        if rep == 5:
            cpu_build = get_tree(t)
        for dst, a in zip(h, B):
            dst[:] = a
        t2 = time.perf_counter()
        t.update()
        t.doBuffersNeedResize()
        t3 = time.perf_counter()
        if rep:
            tb.append(t1 - t0)
            tu.append(t3 - t2)
    cpu_update = get_tree(t)
    out = {"build_" + k: v for k, v in cpu_build.items()}
    out.update({"update_" + k: v for k, v in cpu_update.items()})
    np.savez(os.path.join(d, "out.npz"), **out)
    print(json.dumps({"build_s": float(np.median(tb)), "update_s": float(np.median(tu)), "threads": num_threads(),
                      "total_s": float(sum(tb) + sum(tu))}))


def _counter_rows(path):
    """(kernel name, counter value) per dispatch from a rocprofv3 counter_collection.csv."""
    import csv
    import glob
    files = [path] if os.path.isfile(path) else glob.glob(os.path.join(path, "**", "*counter_collection.csv"),
                                                            recursive=True)
    rows = []
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                rows.append((r.get("Kernel_Name", ""), float(r.get("Counter_Value", "nan"))))
    return rows


def traffic_from_counters(fetch_csv, write_csv, copy_bytes):
    """HBM bytes per launch of the dominant kernel from FETCH_SIZE / WRITE_SIZE runs (KiB per
    dispatch), each scaled by the same counter's reading for the copy kernel of copy_bytes."""
    out = {"kernel": DOMINANT_KERNEL}
    for label, path in (("fetch", fetch_csv), ("write", write_csv)):
        rows = _counter_rows(path)
        copy = [v for k, v in rows if "copyKernel" in k]
        scat = [v for k, v in rows if DOMINANT_KERNEL in k]
        if not copy or not scat:
            raise SystemExit("no copyKernel / %s dispatches in %s" % (DOMINANT_KERNEL, path))
        factor = copy_bytes / (float(np.median(copy)) * 1024.0)
        out[label + "_kib_per_dispatch"] = float(np.median(scat))
        out[label + "_copy_kib"] = float(np.median(copy))
        out[label + "_scale"] = factor
        out[label + "_bytes"] = float(np.median(scat)) * 1024.0 * factor
    out["bytes_per_launch"] = int(out["fetch_bytes"] + out["write_bytes"])
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 25)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (and the oracle checks)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run the known-size copy kernel (for hardware counter scaling)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain measurement")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default sort tuning")
    ap.add_argument("--only-batch", action="store_true", help="run only the batched-vs-plain measurement")
    ap.add_argument("--batch-profile", action="store_true",
                    help="run only the batched path (build + 5 graph-replayed updates, three times) for a kernel "
                         "trace, and print its phase times")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--no-configs", action="store_true", help="skip the other single-tree workloads")
    ap.add_argument("--serve-workers", type=int, default=4, help="worker streams of the serving loop")
    ap.add_argument("--serve-hw-queues", type=int, default=8, choices=range(1, 33), metavar="1..32",
                    help="hardware queues (GPU_MAX_HW_QUEUES) of the serving leg's process")
    ap.add_argument("--serve-rates", default="50,100,200,500,1000,2000", help="arrival rates (requests/s)")
    ap.add_argument("--serve-requests", type=int, default=1000, help="requests per arrival rate")
    ap.add_argument("--serve-saturated", type=int, default=800,
                    help="requests of the saturated trace (all submitted at once; 0 skips it)")
    ap.add_argument("--only-configs", action="store_true", help="run only the other single-tree workloads")
    ap.add_argument("--only-serve", action="store_true", help="run only the serving-loop measurement")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/traffic.json from two rocprofv3 --pmc runs and exit")
    ap.add_argument("--cpu-child", metavar="DIR", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_child:
        cpu_child(args.cpu_child)
        return
    if args.traffic:
        t = traffic_from_counters(args.traffic[0], args.traffic[1], COPY_BYTES)
        with open(os.path.join(ROOT, "profiles", "traffic.json"), "w") as f:
            json.dump(t, f, indent=1)
        print(json.dumps(t))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (requested %d)" % max(args.gpus, world),
              file=sys.stderr)
        sys.exit(2)
    rates = tuple(int(r) for r in args.serve_rates.split(","))
    if args.batch_profile:
        print(json.dumps(batch_profile()))
        return
    if args.only_batch:
        print(json.dumps({"batched": batch_benchmark()}))
        return
    if args.only_serve:
        print(json.dumps({"serving": serving_benchmark(rates, args.serve_requests, workers=args.serve_workers,
                                                       saturated=args.serve_saturated)}))
        return
    if args.only_configs:
        print(json.dumps({"other_configs": other_configs(not args.no_cpu)}))
        return

    out, (A, B, box, after_build, after_update) = headline(args)
    progress("headline done")
    if args.copy_calibration:
        stream = Stream()
        src, dst = DeviceArray(COPY_BYTES), DeviceArray(COPY_BYTES)
        for _ in range(3):
            _lib().call("bench_copy", dst.ptr, src.ptr, COPY_BYTES, stream.p)
        stream.sync()
    if not args.no_configs:
        out["other_configs"] = other_configs(not args.no_cpu)
    if not args.no_batch:
        out["batched"] = batch_benchmark()
        progress("batch done")
    if not args.no_serve:
        out["serving"] = serving_in_child(rates, args.serve_requests, args.serve_workers, args.serve_saturated,
                                          args.serve_hw_queues)
    if not args.no_cpu:
        base, exact_build, exact_update = cpu_baseline(A, B, box, after_build, after_update)
        out["cpu_baseline"] = base
    else:
        out["cpu_baseline"] = None
        exact_build = exact_update = None
    oc = out.get("other_configs", {})
    out["parity"] = {
        "C3_32M_binary_build_vs_oracle": exact_build,
        "C3_32M_binary_moved_graph_update_vs_oracle": exact_update,
        "C1_65k_octree_build_and_views_vs_oracle": oc.get("octree_65k_fp64_uniform", {}).get("bit_exact_vs_oracle"),
        "C2_16M_nfcount_build_vs_oracle": oc.get("octree_16M_fp32_plummer_hilbert_nf4096", {}).get(
            "bit_exact_build_vs_oracle"),
        "C4_512_members_vs_plain_trees": out.get("batched", {}).get("bit_exact_all_members_vs_plain"),
        "full_size_tests": "tests/test_gpu_fullsize.py: C1 build + 3 updates, C2 build + 10 updates + views, "
                           "C3 graph updates, C4 all members vs plain + 32 vs oracle, C5 served mix to 8M",
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
