"""Convert the reference's verification fixtures into compact numpy archives.

Source (MIT licence, nrottstegge/Adaptive_Trees): tests/verification_data/
  snapshot_0000.dat, snapshot_1000.dat  -- 114,537 particles, "charge x y z"
  reference_output/*.csv               -- arrays of Octree<double> after
      build(snapshot_0000) + update(snapshot_1000) with NFCount,
      maxParticlesPerLeaf = 64*64*27, Morton, box [0,1]^3, all views
      (tests/tree_test.cu:155-209).

Output: tests/golden/reference_fixture.npz, and tests/golden/
snapshot_0000_head.dat = the first HEAD_ROWS particle rows of
snapshot_0000.dat with its two '#' header lines, byte for byte (a text
input for the particle-file readers). Coordinates are parsed with
Python's correctly rounded float(), the same value strtod() gives the
reference reader (tests/particle_io.hpp:98-110). leafCounts_d.csv is not
converted: the reference test does not compare it and its content is stale.

Run in a checkout that has the reference at REF (default /root/reference).
"""
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(REF, "tests", "verification_data")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_fixture.npz")
HEAD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "snapshot_0000_head.dat")
HEAD_ROWS = 1500

ARRAYS = {
    "particleKeys_d": np.uint64, "perm_d": np.uint32, "cornerstone_d": np.uint64,
    "sfcBoxIndex_d": np.uint64, "particleBeginIndex_d": np.uint32,
    "particleCounts_d": np.uint32, "hasBoxSplit_d": np.uint8, "boxDepth_d": np.uint8,
    "parentIndex_d": np.int32, "childIndex_d": np.int32,
    "sfcParticleIndex_d": np.uint64, "levelOffset_d": np.int32,
}


def read_snapshot(path):
    xs, ys, zs = [], [], []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            if s.startswith("ESC"):
                e = np.finfo(np.float64).max
                xs.append(e); ys.append(e); zs.append(e)
                continue
            cols = s.split()
            xs.append(float(cols[1])); ys.append(float(cols[2])); zs.append(float(cols[3]))
    return np.array(xs), np.array(ys), np.array(zs)


def read_column(name, dtype):
    with open(os.path.join(SRC, "reference_output", name + ".csv")) as f:
        header = f.readline()
        assert header.strip() == "value", header
        vals = [int(v) for v in f.read().split()]
    if np.issubdtype(dtype, np.signedinteger):
        return np.array(vals, dtype=np.int64).astype(dtype)
    return np.array(vals, dtype=np.uint64).astype(dtype)


def main():
    out = {}
    for tag in ("0000", "1000"):
        x, y, z = read_snapshot(os.path.join(SRC, "snapshot_%s.dat" % tag))
        out["x" + tag], out["y" + tag], out["z" + tag] = x, y, z
    for name, dt in ARRAYS.items():
        out[name] = read_column(name, dt)
    with open(os.path.join(SRC, "reference_output", "scalars.csv")) as f:
        f.readline()
        for line in f:
            k, v = line.strip().split(",")
            out["scalar_" + k] = np.array(int(v), dtype=np.int64)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
    with open(os.path.join(SRC, "snapshot_0000.dat"), "rb") as f:
        lines = f.read().split(b"\n")
    header = [l for l in lines[:2] if l.startswith(b"#")]
    assert len(header) == 2, "expected two '#' header lines"
    with open(HEAD, "wb") as f:
        f.write(b"\n".join(lines[:2 + HEAD_ROWS]) + b"\n")
    print("wrote", HEAD, os.path.getsize(HEAD), "bytes")


if __name__ == "__main__":
    main()
