"""The benchmark-harness driver (adaptive_trees_amd/harness/bench_driver,
the protocol and CSV schema of the reference's benchmarks/drivers/
common.hpp:182-510 with driver_v3.cu) on a small dataset written in the
harness formats: a text first frame with ESC rows, then binary frames with
NaN rows.

Each row's structure columns must equal those of the oracle tree driven
through the same protocol (build frame 0; per later frame: update, and
while buffers need resizing, resizeBuffers + update), recounted from the
oracle's P and K; num_nodes equals lib_num_nodes (a full octree has
L + (L-1)/7 nodes, a binary tree 2L-1). Timing columns are positive
measurements ("nan" on the frame-0 row). Both the graph and the direct path
(BENCH_V3_PATH=direct) and the cold mode are run.
"""
import csv
import os
import subprocess

import numpy as np
import pytest

from adaptive_trees_amd import BINARY, LEAF_COUNT, MORTON, NF_COUNT, OCTREE, make_config
from oracle.binding import OracleTree

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "adaptive_trees_amd", "harness", "bench_driver")
STEP_HEADER = ("version,commit,path,config,dataset,run,snapshot,frame_id,tree_update_ms,view_ms,total_ms,wall_ms,"
               "attempts,maintenance_ms,iterations,num_nodes,num_leaves,max_depth,max_depth_octree_equiv,"
               "num_empty_leaves,empty_leaf_pct,avg_particles_per_leaf,particles_active,particles_escaped,"
               "particle_slots,lib_num_nodes")
COLD_HEADER = ("version,commit,path,config,dataset,process_rep,initial_build_ms,initial_tree_gpu_ms,"
               "initial_view_gpu_ms,num_leaves,num_nodes,max_depth")
CONFIGS = {"octree_leafcount": (OCTREE, LEAF_COUNT, 64, 3), "octree_nfcount": (OCTREE, NF_COUNT, 64 * 64 * 27, 3),
           "kdtree3d_leafcount": (BINARY, LEAF_COUNT, 64, 1)}
BOX = (-0.25, 1.25, -0.5, 1.5, 0.0, 1.0)


@pytest.fixture(scope="module")
def dataset(tmp_path_factory):
    d = tmp_path_factory.mktemp("halo")
    rng = np.random.default_rng(11)
    n = 24000
    pts = np.concatenate([0.5 + 0.02 * rng.standard_normal((n // 2, 3)), rng.random((n - n // 2, 3))]).clip(0, 1)
    esc = np.zeros(n, bool)
    esc[[5, 777, 9001]] = True
    frames = []
    with open(d / "frame0.dat", "w") as f:
        f.write("# x y z charge\n")
        for i in range(n):
            f.write("ESC\n" if esc[i] else "%.17g %.17g %.17g 1.0\n" % tuple(pts[i]))
    frames.append(pts.copy())
    (d / "bin").mkdir()
    for k in range(1, 4):
        pts = (pts + rng.normal(0, 0.004 * k, pts.shape)).clip(0, 1)
        if k == 2:
            esc[[100, 200]] = True  # two more slots escape
        v = pts.copy()
        v[esc] = np.nan
        v.astype(np.float64).tofile(d / "bin" / ("noisy_%03d.bin" % k))
        frames.append(pts.copy())
    with open(d / "manifest.txt", "w") as f:
        f.write("dataset halo\nslots %d\nbox %s\n" % (n, " ".join("%.17g" % b for b in BOX)))
        f.write("frame 0 0 xyzc frame0.dat\n")
        for k in range(1, 4):
            f.write("frame %d %d bin bin/noisy_%03d.bin\n" % (k, 10 * k, k))
    masks = []
    e = np.zeros(n, bool)
    e[[5, 777, 9001]] = True
    for k in range(4):
        if k == 2:
            e[[100, 200]] = True
        masks.append(e.copy())
    return str(d / "manifest.txt"), frames, masks


def run(args, env_path=None):
    assert os.path.exists(DRIVER), "adaptive_trees_amd/harness/bench_driver is not built"
    env = dict(os.environ)
    env.pop("BENCH_V3_PATH", None)
    if env_path:
        env["BENCH_V3_PATH"] = env_path
    return subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=300, env=env)


def oracle_rows(cfg_name, frames, masks):
    kind, crit, limit, bits = CONFIGS[cfg_name]
    big = np.finfo(np.float64).max
    coords = [np.where(masks[0], big, frames[0][:, a]).copy() for a in range(3)]
    n = len(coords[0])
    cfg = make_config(max_particles_per_leaf=limit, split_criterion=crit, sfc_kind=MORTON, box=BOX,
                      views=(True, True, True))
    t = OracleTree(*coords, n, cfg, real=np.float64, kind=kind)
    t.build()
    t.computeViews()
    while t.doBuffersNeedResize():
        t.resizeBuffers()
        t.computeViews()
    out = [summarise(t, bits)]
    for k in range(1, len(frames)):
        for a in range(3):
            coords[a][:] = np.where(masks[k], big, frames[k][:, a])
        attempts = 0
        while True:
            t.update()
            t.computeViews()
            attempts += 1
            if not t.doBuffersNeedResize():
                break
            t.resizeBuffers()
        out.append(summarise(t, bits) + (attempts,))
    return out


def summarise(t, bits):
    P = t.get("particleKeys_d").astype(np.uint64)
    K = t.get("cornerstone_d").astype(np.uint64)
    pos = np.searchsorted(P, K, side="left")
    leaves = len(K) - 1
    ranges = (K[1:] - K[:-1]).astype(np.uint64)
    range_bits = np.array([int(r).bit_length() - 1 for r in ranges])
    depth = int(((63 - range_bits) // bits).max())
    active = int(pos[-1] - pos[0])
    empty = int((pos[1:] == pos[:-1]).sum())
    nodes = leaves + (leaves - 1) // ((1 << bits) - 1)
    return (leaves, nodes, depth, empty, active, len(P) - active, len(P), t.numNodes())


@pytest.mark.parametrize("path", ["graph", "direct"])
@pytest.mark.parametrize("cfg_name", list(CONFIGS))
def test_steps_rows_match_the_oracle(tmp_path, dataset, cfg_name, path):
    manifest, frames, masks = dataset
    out = str(tmp_path / "steps.csv")
    print("harness", cfg_name, path, flush=True)
    r = run(["--manifest", manifest, "--config", cfg_name, "--mode", "steps", "--version", "amd", "--commit", "abc",
             "--run", "2", "--out", out], "direct" if path == "direct" else None)
    assert r.returncode == 0, r.stderr
    with open(out) as f:
        lines = f.read().splitlines()
    assert lines[0] == STEP_HEADER
    rows = list(csv.DictReader(lines))
    assert len(rows) == 4
    want = oracle_rows(cfg_name, frames, masks)
    for k, (row, w) in enumerate(zip(rows, want)):
        assert (row["version"], row["commit"], row["path"], row["config"], row["dataset"], row["run"]) == \
            ("amd", "abc", path, cfg_name, "halo", "2")
        assert int(row["snapshot"]) == k and int(row["frame_id"]) == 10 * k
        got = (int(row["num_leaves"]), int(row["num_nodes"]), int(row["max_depth"]), int(row["num_empty_leaves"]),
               int(row["particles_active"]), int(row["particles_escaped"]), int(row["particle_slots"]),
               int(float(row["lib_num_nodes"])))
        assert got == w[:8], (k, got, w)
        assert int(row["num_nodes"]) == int(float(row["lib_num_nodes"]))
        assert row["iterations"] == "-1"
        if k == 0:
            assert row["tree_update_ms"] == "nan" and row["attempts"] == "0"
        else:
            assert int(row["attempts"]) == w[8]
            for c in ("tree_update_ms", "view_ms", "total_ms", "wall_ms"):
                assert float(row[c]) > 0, (c, row[c])
            assert float(row["maintenance_ms"]) >= 0
        bits = CONFIGS[cfg_name][3]
        equiv = float(row["max_depth_octree_equiv"])
        assert equiv == pytest.approx(w[2] / 3.0 if bits == 1 else w[2], rel=1e-9)


def test_cold_mode_and_errors(tmp_path, dataset):
    manifest, frames, masks = dataset
    out = str(tmp_path / "cold.csv")
    r = run(["--manifest", manifest, "--config", "octree_leafcount", "--mode", "cold", "--version", "v",
             "--commit", "c", "--run", "0", "--out", out])
    assert r.returncode == 0, r.stderr
    with open(out) as f:
        lines = f.read().splitlines()
    assert lines[0] == COLD_HEADER and len(lines) == 2
    row = next(csv.DictReader(lines))
    w = oracle_rows("octree_leafcount", frames[:1], masks[:1])[0]
    assert (int(row["num_leaves"]), int(row["num_nodes"]), int(row["max_depth"])) == (w[0], w[1], w[2])
    assert float(row["initial_build_ms"]) > 0 and float(row["initial_tree_gpu_ms"]) > 0
    # usage and input errors end with status 1 and a message
    assert run(["--config", "octree_leafcount"]).returncode == 1
    bad = run(["--manifest", str(tmp_path / "none.txt"), "--config", "octree_leafcount", "--out", out])
    assert bad.returncode == 1 and "manifest" in bad.stderr
    assert run(["--manifest", manifest, "--config", "quadtree", "--out", out]).returncode == 1
