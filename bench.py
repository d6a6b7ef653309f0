"""Benchmark of the adaptive-tree build/update path on one MI355X.

Workload (the 32M-particle build the headline roofline target is quoted on):
binary longest-side tree, N = 2^25 FP32 particles in a 4:1:1 box, 40% + 40%
isotropic Gaussian clusters (sigma = 0.02 of the long side, centres at 1/4
and 3/4 of x) + 20% uniform background, LeafCount 32.

One step = build() followed by one hipGraph-replayed update() of the same
resident particles; value = 2 N / step time (particles through the path per
second). Coordinates are generated on the host, uploaded once, and stay in
HBM for the whole timed region.

Output: one JSON line (metric, value, roofline of the dominant kernel, CPU
baseline of the oracle on a bounded sample, per-phase timings).
"""
import argparse
import ctypes
import json
import os
import re
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


def algorithmic_build_bytes(n, leaves, s=4):
    """Bytes any build must move: key generation reads 3s and writes 8 per
    particle, a sort reads the keys and writes the sorted (key, slot) pairs
    (8 + 12), the construction reads P once and writes K and N (8 + 4 per
    leaf). A floor for the whole build, not a count of what this one moves."""
    return n * (3 * s + 8) + n * 20 + 8 * n + 12 * leaves


def cpu_baseline(x, y, z, x1, y1, z1, box, min_seconds):
    """The oracle (oracle/, OpenMP C++) on the same workload in host memory: build() then one
    update() on the moved positions, repeated until min_seconds of CPU work; the median repeat
    is reported."""
    from oracle.binding import OracleTree, num_threads
    n = len(x)
    hx, hy, hz = x.copy(), y.copy(), z.copy()
    cfg = make_config(max_particles_per_leaf=32, box=box)
    t = OracleTree(hx, hy, hz, n, cfg, real=np.float32, kind=BINARY)
    times = []
    start = time.perf_counter()
    while len(times) < 3 or (time.perf_counter() - start < min_seconds and len(times) < 50):
        hx[:], hy[:], hz[:] = x, y, z
        t0 = time.perf_counter()
        t.build()
        hx[:], hy[:], hz[:] = x1, y1, z1
        t.update()
        t.doBuffersNeedResize()
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    return {"value": round(2 * n / dt / 1e6, 3), "unit": "Mparticles/s", "cores": num_threads(),
            "kind": "port",
            "sample": "oracle build() + one update() of the full %d-particle workload in host memory, "
                      "median of %d repeats (%.2f s each, %.1f s total)" % (n, len(times), dt, sum(times))}


def batch_workload(trees, seed):
    """512 octree inputs of the batched workload: sizes log-uniform in [1e3, 2e6], uniform or
    Plummer per tree by a seed bit, FP32, plus one reflected-Gaussian moved snapshot
    (sigma = 1e-3 of the unit box). Returns sizes and two concatenated (x, y, z) snapshots."""
    sizes = workloads.batch_sizes(trees, seed=seed)
    rng = np.random.default_rng(seed)
    kinds = rng.integers(0, 2, size=trees)
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
    the same in both paths)."""
    lib = _lib()
    sizes, first, moved = batch_workload(trees, seed)
    progress("batched workload generated: %d trees, %d particles" % (trees, int(sizes.sum())))
    total = int(sizes.sum())
    off = np.concatenate([[0], np.cumsum(sizes)])
    cur = [DeviceArray(a) for a in first]
    snaps = [[DeviceArray(a) for a in first], [DeviceArray(a) for a in moved]]
    nbytes = total * 4
    stream = Stream()
    cfg = dict(max_particles_per_leaf=64)

    def ptrs(a):
        return [cur[a].ptr + 4 * int(off[k]) for k in range(trees)]

    def load(r):
        for a in range(3):
            lib.call("memcpy_async", cur[a].ptr, snaps[r % 2][a].ptr, nbytes, 3, stream.p)

    def resize_plain(ts):
        for t in ts:
            while t.doBuffersNeedResize():
                t.resizeBuffers()
                t.update()

    # plain: 512 independent trees on one stream
    plain = [Tree(*(ptrs(a)[k] for a in range(3)), int(sizes[k]), make_config(stream=stream.p, **cfg),
                  real=np.float32, kind=OCTREE) for k in range(trees)]

    def run_plain():
        load(0)
        for t in plain:
            t.build()
        for r in range(1, rounds + 1):
            load(r)
            for t in plain:
                t.update()
            resize_plain(plain)
        stream.sync()

    batch = Batch(ptrs(0), ptrs(1), ptrs(2), sizes, make_config(stream=stream.p, **cfg), real=np.float32)
    graph = [None]

    def run_batch():
        load(0)
        batch.build()
        graph[0] = graph[0] or batch.capture_update(stream.p)
        for r in range(1, rounds + 1):
            load(r)
            lib.call("graph_launch", graph[0], stream.p)
            while batch.needsResize().any():
                batch.resizeBuffers()
                lib.call("graph_destroy", graph[0])
                graph[0] = batch.capture_update(stream.p)
                lib.call("graph_launch", graph[0], stream.p)
        stream.sync()

    run_plain()  # warm-up: allocations reach their steady state
    run_batch()
    progress("batched/plain warm-up done")
    t0 = time.perf_counter()
    run_plain()
    plain_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    run_batch()
    batch_s = time.perf_counter() - t0
    sample = list(range(0, trees, max(1, trees // 32)))
    exact = all(np.array_equal(batch.tree(k).get(n), plain[k].get(n))
                for k in sample for n in ("particleKeys_d", "perm_d", "cornerstone_d", "leafCounts_d"))
    work = total * (rounds + 1)
    return {
        "workload": "%d octrees, sizes log-uniform [1e3, 2e6] (total %d), FP32 uniform/Plummer, Morton, "
                    "LeafCount 64; build + %d graph-replayed updates (sigma 1e-3 reflected steps)"
                    % (trees, total, rounds),
        "particles": total,
        "plain_s": round(plain_s, 4),
        "batched_s": round(batch_s, 4),
        "speedup": round(plain_s / batch_s, 3),
        "batched_Mparticles_per_s": round(work / batch_s / 1e6, 1),
        "plain_Mparticles_per_s": round(work / plain_s / 1e6, 1),
        "bit_exact_sample_trees": len(sample),
        "bit_exact": bool(exact),
    }


def other_configs():
    """The other single-tree workloads: 65,536 FP64 uniform octree (Morton, LeafCount 64) build,
    checked bit for bit against the oracle; and a 2^24-particle FP32 Plummer octree (a = 0.05,
    Hilbert, NFCount 4096) build followed by 10 updates of reflected Gaussian steps (sigma 1e-3)."""
    from oracle.binding import OracleTree
    lib = _lib()
    stream = Stream()
    e0, e1 = Event(), Event()
    out = {}
    # 65k FP64 uniform octree
    n = 65536
    x, y, z = workloads.uniform_cube(n, seed=1, dtype=np.float64)
    dev = [DeviceArray(a) for a in (x, y, z)]
    t = Tree(*dev, n, make_config(max_particles_per_leaf=64, stream=stream.p), real=np.float64, kind=OCTREE)
    times = []
    for _ in range(25):
        e0.record(stream)
        t.build()
        e1.record(stream)
        stream.sync()
        times.append(e1.ms_since(e0))
    o = OracleTree(x.copy(), y.copy(), z.copy(), n, make_config(max_particles_per_leaf=64), real=np.float64,
                   kind=OCTREE)
    o.build()
    exact = all(np.array_equal(t.get(k), o.get(k)) for k in ("particleKeys_d", "perm_d", "cornerstone_d",
                                                              "leafCounts_d"))
    out["octree_65k_fp64_uniform_build"] = {"build_ms_median": round(float(np.median(times[5:])), 4),
                                            "leaves": int(t.scalars().num_leaves), "bit_exact_vs_oracle": exact}
    t.close()
    progress("65k octree done")
    # 16M FP32 Plummer Hilbert NFCount octree, build + 10 updates
    n = 1 << 24
    pts = workloads.plummer(n, seed=2, dtype=np.float32)
    dev = [DeviceArray(a) for a in pts]
    t = Tree(*dev, n, make_config(max_particles_per_leaf=4096, split_criterion=NF_COUNT, sfc_kind=HILBERT,
                                  stream=stream.p), real=np.float32, kind=OCTREE)
    e0.record(stream)
    t.build()
    e1.record(stream)
    stream.sync()
    build_ms = e1.ms_since(e0)
    upd, resizes, cur = [], 0, pts
    for step in range(10):
        cur = workloads.reflect_step(cur, 1e-3, seed=200 + step)
        for d, a in zip(dev, cur):
            d.upload(a)
        e0.record(stream)
        t.update()
        e1.record(stream)
        stream.sync()
        upd.append(e1.ms_since(e0))
        while t.doBuffersNeedResize():
            resizes += 1
            t.resizeBuffers()
            t.update()
    out["octree_16M_fp32_plummer_hilbert_nf4096"] = {
        "build_ms": round(build_ms, 3), "update_ms_mean": round(float(np.mean(upd)), 3),
        "update_ms_max": round(float(np.max(upd)), 3), "resizes": resizes,
        "leaves": int(t.scalars().num_leaves),
        "Mparticles_per_s": round(11 * n / ((build_ms + sum(upd)) * 1e-3) / 1e6, 1)}
    t.close()
    progress("16M NF octree done")
    return out


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


def serving_benchmark(rates=(50, 200, 1000, 2000, 10000), requests=160, seed=5, workers=4):
    """The serving loop (4 worker streams, 200 GB budget) against serial plain calls on the
    same trace; latency = result ready - arrival, per request."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(1e4), np.log(8e6), size=20))).astype(np.int64)
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

    results = []
    for rate in rates:
        trace = serving_trace(rate, requests, pool, seed + rate)
        # inputs of every request, prepared before the clock starts
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
        info = {}
        for r in trace:
            if r[0] == "build":
                info[r[2]] = r
        # served
        server = Server(hbm_budget_bytes=200 << 30, workers=workers, max_batch=16)
        ids, turn, reqs = {}, {}, []
        start = time.perf_counter()
        for r, bufs in zip(trace, prep):
            delay = r[1] - (time.perf_counter() - start)
            if delay > 0:
                time.sleep(delay)
            if r[0] == "build":
                _, _, tid, p, kind, real, cfg = r
                ids[tid], rid = server.submit_build(*bufs[0], len(pool[p][0][0]), make_config(**cfg), real=real,
                                                    kind=kind)
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
        lat, first, last, work = [], None, 0.0, 0
        for rid, n in reqs:
            assert server.wait(rid) == 0
            s_ms, b_ms, d_ms = server.times(rid)
            lat.append(d_ms - s_ms)
            first = s_ms if first is None else min(first, s_ms)
            last = max(last, d_ms)
            work += n
        stats = server.stats()
        server.close()
        served_s = (last - first) / 1e3
        # serial plain calls: one Tree per build, requests in arrival order on one stream
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
        results.append({
            "rate_per_s": rate, "requests": len(reqs),
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "serial_p50_ms": round(float(np.percentile(slat, 50)), 3),
            "serial_p99_ms": round(float(np.percentile(slat, 99)), 3),
            "requests_per_s": round(len(reqs) / served_s, 1),
            "Mparticles_per_s": round(work / served_s / 1e6, 1),
            "serial_requests_per_s": round(len(reqs) / serial_s, 1),
            "batched_updates": int(stats["batched_updates"]), "peak_GB": round(stats["bytes_peak"] / 1e9, 3),
        })
        progress("serving rate %d/s done" % rate)
    return {"workload": "Poisson trace, %d requests per rate, sizes log-uniform [1e4, 8e6], 70%% octree / 30%% "
                        "binary, FP32/FP64, Hilbert/Morton, LeafCount 64 / NFCount 4096; %d worker streams with "
                        "per-tree update graphs, 200 GB budget; serial = plain Tree calls in arrival order; "
                        "10000/s is an overload probe of capacity" % (requests, workers),
            "rates": results}


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
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU-baseline work to time")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run the known-size copy kernel (for hardware counter scaling)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain measurement")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default sort tuning")
    ap.add_argument("--only-batch", action="store_true", help="run only the batched-vs-plain measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--no-configs", action="store_true", help="skip the other single-tree workloads")
    ap.add_argument("--serve-workers", type=int, default=4, help="worker streams of the serving loop")
    ap.add_argument("--serve-rates", default="50,200,1000,2000,10000", help="arrival rates (requests/s)")
    ap.add_argument("--only-configs", action="store_true", help="run only the other single-tree workloads")
    ap.add_argument("--only-serve", action="store_true", help="run only the serving-loop measurement")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/traffic.json from two rocprofv3 --pmc runs and exit")
    args = ap.parse_args()
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

    if args.only_batch:
        print(json.dumps({"batched": batch_benchmark()}))
        return
    if args.only_serve:
        rates = tuple(int(r) for r in args.serve_rates.split(","))
        print(json.dumps({"serving": serving_benchmark(rates=rates, workers=args.serve_workers)}))
        return
    if args.only_configs:
        print(json.dumps({"other_configs": other_configs()}))
        return

    n = args.n
    box = workloads.ANISO_BOX
    tuning = tune.get() if args.no_autotune else tune.autotune(n)
    progress("sort tuning: %s" % json.dumps(tuning))
    x, y, z = workloads.two_clusters(n, seed=3, dtype=np.float32)
    dx, dy, dz = DeviceArray(x), DeviceArray(y), DeviceArray(z)
    stream = Stream()
    cfg = make_config(max_particles_per_leaf=32, box=box, stream=stream.p)
    tree = Tree(dx, dy, dz, n, cfg, real=np.float32, kind=BINARY)
    tree.build()
    graph = ctypes.c_void_p()
    _lib().call("graph_capture_update", tree.handle, stream.p, ctypes.byref(graph))
    addr = tree.buffer_info("cornerstone_d")[0]

    def step():
        tree.build()
        _lib().call("graph_launch", graph.value, stream.p)

    for _ in range(args.warmup):
        step()
    stream.sync()
    assert tree.buffer_info("cornerstone_d")[0] == addr, "build moved the captured buffers"

    # phase timings (outside the timed region)
    e0, e1, e2 = Event(), Event(), Event()
    tree.build()
    e0.record(stream)
    _lib().call("graph_launch", graph.value, stream.p)
    e1.record(stream)
    stream.sync()
    update_ms = e1.ms_since(e0)
    t0 = time.perf_counter()
    tree.build()
    stream.sync()
    build_wall_ms = (time.perf_counter() - t0) * 1e3

    _lib().call("profile_reset")
    _lib().call("profile_enable", 1)
    stream.sync()
    e0.record(stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    e2.record(stream)
    stream.sync()
    wall = time.perf_counter() - t0
    _lib().call("profile_enable", 0)
    ms_per_step = wall * 1e3 / args.steps
    event_ms = e2.ms_since(e0) / args.steps

    leaves = int(tree.scalars().num_leaves)
    def per_launch(label):
        ms, cnt = profile_query(label)
        return ms / max(cnt, 1), cnt

    dom_ms, dom_n = per_launch(DOMINANT_LABEL)
    key_ms, _ = per_launch("keygen")
    pass_ms, _ = per_launch("sort_pass")
    level_ms, level_n = per_launch("msd_level")
    dir_ms, dir_n = profile_query("bin_direct")
    achieved = PAIR_BYTES * n / (dom_ms * 1e-3) / 1e9 if dom_n else 0.0
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            t = json.load(f)
            if t.get("kernel") == DOMINANT_KERNEL:
                traffic = t.get("bytes_per_launch")

    # an update that rebalances: reflected Gaussian steps (sigma = 1e-3 of the box)
    x1, y1, z1 = workloads.reflect_step((x, y, z), 1e-3, seed=33, box=box)
    moved_ms = []
    for r in range(3):
        tree.build()
        for d, a in ((dx, x1), (dy, y1), (dz, z1)):
            d.upload(a)
        stream.sync()
        e0.record(stream)
        _lib().call("graph_launch", graph.value, stream.p)
        e1.record(stream)
        stream.sync()
        moved_ms.append(e1.ms_since(e0))
        for d, a in ((dx, x), (dy, y), (dz, z)):
            d.upload(a)
    tree.doBuffersNeedResize()

    if args.copy_calibration:
        nbytes = COPY_BYTES
        src, dst = DeviceArray(nbytes), DeviceArray(nbytes)
        for _ in range(3):
            _lib().call("bench_copy", dst.ptr, src.ptr, nbytes, stream.p)
        stream.sync()

    build_bytes = algorithmic_build_bytes(n, leaves)
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
            "workload": "binary longest-side tree build() + graph-replayed update(), N=2^25 FP32, "
                        "4:1:1 box, two Gaussian clusters + 20% uniform, LeafCount 32",
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
            "build_wall": round(build_wall_ms, 4),
            "update_static_graph": round(update_ms, 4),
            "update_moved_graph": round(float(np.median(moved_ms)), 4),
            "keygen": round(key_ms, 4),
            "sort_level1_pass": round(pass_ms, 4),
            "sort_levels_2_to_8": round(level_ms * level_n / max(dom_n, 1), 4),
            "sort_lds_chunks": round(dom_ms, 4),
            "binary_construction": round(dir_ms / max(dir_n, 1) * 2, 4),
        },
        "build_roofline": {
            "algorithmic_bytes": int(build_bytes),
            "frac_of_peak": round(build_bytes / (build_wall_ms * 1e-3) / (PEAK_HBM_GBS * 1e9), 4),
        },
    }
    tree.close()
    if not args.no_configs:
        out["other_configs"] = other_configs()
    if not args.no_batch:
        out["batched"] = batch_benchmark()
    if not args.no_serve:
        out["serving"] = serving_benchmark(workers=args.serve_workers)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(x, y, z, x1, y1, z1, box, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
