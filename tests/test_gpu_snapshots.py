"""The snapshot benchmark (adaptive_trees_amd/harness/bench_snapshots: the
reference's tests/bench_snapshots.cu protocol) on small datasets.

The program runs every update through a directly enqueued tree and a
graph-replayed tree, requires both to be equal after every comparison (it
exits 1 otherwise) and writes one CSV row per snapshot. Each row's structure
columns must equal those of the oracle tree driven through the same
protocol: build on the first frame with the program's padded bounding box,
all views (with the resize protocol), then per comparison update + views,
repeated after resizeBuffers while buffers need resizing. Cases: a single
"x y z charge" file (one row: the mean of --repeats warm comparisons) and a
folder of "charge x y z" snapshot_NNNNNN.dat frames with escaped (ESC) rows
and a stride, for the three runs (binary LeafCount, octree LeafCount and
NFCount).
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
PROGRAM = os.path.join(ROOT, "adaptive_trees_amd", "harness", "bench_snapshots")
HEADER = ("snapshot,h2d_ms,init_tree_ms,tree_ms,view_ms,compute_total_ms,total_with_h2d_ms,num_nodes,num_leaves,"
          "num_empty_leaves,empty_leaf_percent,avg_particles_per_leaf,max_depth,graph_tree_ms,graph_view_ms,"
          "graph_compute_total_ms,maintenance_ms,graph_maintenance_ms,graph_initial_capture_ms,num_particle_slots,"
          "num_active_particles,num_escaped_particles,tree_type,split_criterion")
BIG = np.finfo(np.float64).max
RUNS = [("binary", "leafcount", BINARY, LEAF_COUNT, 32), ("octree", "leafcount", OCTREE, LEAF_COUNT, 32),
        ("octree", "nfcount", OCTREE, NF_COUNT, 2000)]


def padded_box(frames):
    """The program's box: bounds of the active particles of all frames, each
    axis widened by max(5% of its extent, 8 eps max(1, |lo|, |hi|))."""
    pts = np.concatenate([f[~np.all(f == BIG, axis=1)] for f in frames])
    box = []
    for a in range(3):
        lo, hi = pts[:, a].min(), pts[:, a].max()
        m = max(0.05 * (hi - lo), 8 * np.finfo(np.float64).eps * max(1.0, abs(lo), abs(hi)))
        box += [lo - m, hi + m]
    return tuple(box)


def stats(t, kind):
    K = t.get("cornerstone_d").astype(np.uint64)
    N = t.get("leafCounts_d").astype(np.int64)
    L = len(N)
    r = (K[1:] - K[:-1]).astype(np.uint64)
    if kind == BINARY:
        depth = max(64 - int(v).bit_length() for v in r)
    else:
        depth = max((64 - int(v - np.uint64(1)).bit_length() - 1) // 3 if v > 1 else 21 for v in r)
    nodes = L + (L - 1) // (1 if kind == BINARY else 7)
    active = int(N.sum())
    assert active == t.numActiveParticles()
    return dict(num_nodes=nodes, num_leaves=L, num_empty_leaves=int((N == 0).sum()), max_depth=depth,
                num_particle_slots=t.numParticles(), num_active_particles=active,
                num_escaped_particles=t.numParticles() - active)


def oracle_rows(frames, kind, crit, limit, box, comparisons):
    """Rows after comparisons[k] comparisons on frame k (the warm-up
    comparison on the first frame included)."""
    coords = [frames[0][:, a].copy() for a in range(3)]
    cfg = make_config(max_particles_per_leaf=limit, split_criterion=crit, sfc_kind=MORTON, box=box)
    t = OracleTree(*coords, len(coords[0]), cfg, real=np.float64, kind=kind)
    t.build()

    def views():
        t.computeViews((1, 1, 1))
        while t.doBuffersNeedResize():
            t.resizeBuffers()
            t.computeViews((1, 1, 1))

    def compare_once():
        while True:
            t.update()
            t.computeViews((1, 1, 1))
            if not t.doBuffersNeedResize():
                return
            t.resizeBuffers()

    views()
    rows = []
    for k, f in enumerate(frames):
        for a in range(3):
            coords[a][:] = f[:, a]
        for _ in range(comparisons[k]):
            compare_once()
        rows.append(stats(t, kind))
    return rows


def run(args, cwd):
    assert os.path.exists(PROGRAM), "adaptive_trees_amd/harness/bench_snapshots is not built"
    return subprocess.run([PROGRAM] + args, capture_output=True, text=True, timeout=600, cwd=cwd)


def check_rows(rows, want, tree, crit):
    assert len(rows) == len(want)
    for row, w in zip(rows, want):
        assert (row["tree_type"], row["split_criterion"]) == (tree, crit)
        got = {k: int(row[k]) for k in w}
        assert got == w, (got, w)
        for c in ("tree_ms", "view_ms", "compute_total_ms"):
            assert float(row[c]) > 0, (c, row[c])


def cloud(rng, n):
    return np.concatenate([0.3 + 0.05 * rng.standard_normal((n // 2, 3)), rng.random((n - n // 2, 3)) * 2.0 - 0.5])


def test_single_file_repeats(tmp_path):
    rng = np.random.default_rng(21)
    pts = cloud(rng, 30000)
    with open(tmp_path / "halo.dat", "w") as f:
        for p in pts:
            f.write("%.17g %.17g %.17g 1.0\n" % tuple(p))
    repeats = 3
    r = run(["--dataset", str(tmp_path / "halo.dat"), "--repeats", str(repeats), "--leaf-limit", "32",
             "--nf-limit", "2000", "--csv", str(tmp_path / "out.csv")], str(tmp_path))
    assert r.returncode == 0, r.stderr
    lines = open(tmp_path / "out.csv").read().splitlines()
    assert lines[0] == HEADER
    rows = list(csv.DictReader(lines))
    assert len(rows) == 3
    box = padded_box([pts])
    for row, (tree, crit, kind, c, limit) in zip(rows, RUNS):
        want = oracle_rows([pts], kind, c, limit, box, [1 + repeats])
        check_rows([row], want, tree, crit)
        for g in ("graph_tree_ms", "graph_view_ms", "graph_compute_total_ms", "graph_initial_capture_ms"):
            assert float(row[g]) > 0, g


def test_series_with_escapes_and_stride(tmp_path):
    rng = np.random.default_rng(22)
    n = 20000
    pts = cloud(rng, n)
    esc = np.zeros(n, bool)
    d = tmp_path / "series"
    d.mkdir()
    frames = []
    for k in range(0, 7):
        pts = pts + rng.normal(0, 0.01, pts.shape)
        if k == 3:
            esc[rng.choice(n, 50, replace=False)] = True
        if k == 5:
            esc[:10] = False  # reactivation of slots that escaped earlier (if any)
        f = np.where(esc[:, None], BIG, pts)
        with open(d / ("snapshot_%06d.dat" % k), "w") as fh:
            fh.write("# charge x y z\n")
            for i in range(n):
                fh.write("ESC\n" if esc[i] else "1.0 %.17g %.17g %.17g\n" % tuple(pts[i]))
        if k % 2 == 0:
            frames.append(f)
    (d / "notes.txt").write_text("not a frame\n")
    r = run(["--dataset", str(d), "--series", "--stride", "2", "--last-snapshot", "6", "--leaf-limit", "32",
             "--nf-limit", "2000", "--csv", str(tmp_path / "series.csv")], str(tmp_path))
    assert r.returncode == 0, r.stderr
    rows = list(csv.DictReader(open(tmp_path / "series.csv").read().splitlines()))
    assert len(rows) == 3 * len(frames)
    box = padded_box(frames)
    for j, (tree, crit, kind, c, limit) in enumerate(RUNS):
        mine = rows[j * len(frames):(j + 1) * len(frames)]
        assert [int(x["snapshot"]) for x in mine] == [0, 2, 4, 6]
        # every row follows one comparison on its frame (frame 0's is the
        # warm-up comparison that precedes its row)
        want = oracle_rows(frames, kind, c, limit, box, [1, 1, 1, 1])
        check_rows(mine, want, tree, crit)
        assert mine[0]["graph_tree_ms"] == "nan" and float(mine[0]["graph_initial_capture_ms"]) > 0
        for x in mine[1:]:
            assert float(x["graph_tree_ms"]) > 0 and float(x["h2d_ms"]) > 0


def test_usage_errors(tmp_path):
    assert run(["--tree", "binary", "--criterion", "nfcount"], str(tmp_path)).returncode == 1
    assert run(["--dataset", str(tmp_path / "missing.dat")], str(tmp_path)).returncode == 1
    (tmp_path / "mixed").mkdir()
    for name in ("xyz_1", "snapshot_000002.dat"):
        (tmp_path / "mixed" / name).write_text("1 2 3 4\n")
    bad = run(["--dataset", str(tmp_path / "mixed")], str(tmp_path))
    assert bad.returncode == 1 and "Mixed snapshot formats" in bad.stderr
