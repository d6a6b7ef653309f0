"""print() diagnostics (at_tree_print; src/octree/octree_print.cu:57-382,
src/kdtree3d/kdtree3d_print.cu:11-43).

The expected text is assembled here from the oracle tree's getters (the
values come from the CPU restatement, independent of the GPU) in the
reference's layout: section order, labels, sample counts (4 head + 4 tail
rows, 3 + 3 in the one-line arrays), "..." rules, fixed two-decimal
averages and percentages, uint8 arrays printed as numbers, binary leaf
depth = count of leading zeros of the key range, default stream formatting
(%g, six digits) for the binary mean. Octree LeafCount/NFCount with and
without views, and a binary tree, must match exactly. (The NF section of a
LeafCount octree shows whatever its NF buffer holds -- unset in the
reference -- so there the expected text takes the GPU tree's own getter.)
"""
import numpy as np
import pytest

from adaptive_trees_amd import BINARY, HILBERT, MORTON, NF_COUNT, OCTREE, DeviceArray, Tree, make_config
from oracle.binding import OracleTree

pytestmark = pytest.mark.gpu

SEP = "-" * 64


def sample_lines(name, v):
    out = ["%s size = %d" % (name, len(v))]
    if len(v) == 0:
        return out
    head = min(4, len(v))
    out += ["  %s[%d] = %d" % (name, i, int(v[i])) for i in range(head)]
    if len(v) > 8:
        out.append("  ...")
    tail = max(head, len(v) - 4) if len(v) > 4 else head
    out += ["  %s[%d] = %d" % (name, i, int(v[i])) for i in range(tail, len(v))]
    return out


def array_line(name, v):
    s = "%-18s: size=%d" % (name, len(v))
    if len(v):
        s += "  min=%d  max=%d" % (int(np.min(v)), int(np.max(v)))
    items = [str(int(x)) for x in v[:min(3, len(v))]]
    if len(v) > 6:
        items.append("...")
    if len(v) > 3:
        items += [str(int(x)) for x in v[max(min(3, len(v)), len(v) - 3):]]
    return s + "  samples=[" + ", ".join(items) + "]"


def octree_text(t, depth, nf=None):
    P, Perm = t.get("particleKeys_d"), t.get("perm_d")
    K, N = t.get("cornerstone_d"), t.get("leafCounts_d")
    NF = t.get("nearFieldCounts_d") if nf is None else nf
    sc = t.scalars()
    L = max(len(K) - 1, 0)
    o = [SEP, "Adaptive Octree Summary", SEP,
         "Split criterion: " + ("NFCount" if sc.split_criterion == NF_COUNT else "LeafCount"),
         "SFC kind       : " + ("Hilbert" if sc.sfc_kind == HILBERT else "Morton"),
         "Particles      : %d" % sc.num_active_particles, "Input slots    : %d" % sc.num_particles,
         "Leaves         : %d" % L, "", SEP, "Leaf occupancy", SEP]
    if len(N):
        e = int((N == 0).sum())
        o += ["Min particles/leaf : %d" % N.min(), "Max particles/leaf : %d" % N.max(),
              "Avg particles/leaf : %.2f" % (int(N.sum(dtype=np.uint64)) / len(N)),
              "Empty leaves        : %d (%.2f%%)" % (e, 100.0 * e / len(N))]
    o += ["", SEP, "Nearfield", SEP]
    if len(NF):
        e = int((NF == 0).sum())
        tot = int(NF.sum(dtype=np.uint64))
        o += ["Min interactions/leaf : %d" % NF.min(), "Max interactions/leaf : %d" % NF.max(),
              "Total number interactions   : %d" % tot, "Avg interactions/leaf : %.2f" % (tot / len(NF)),
              "Empty leaves           : %d (%.2f%%)" % (e, 100.0 * e / len(NF))]
    o += ["", SEP, "Cornerstone leaf ranges", SEP]
    if len(K) >= 2:
        d = K[1:].astype(np.uint64) - K[:-1].astype(np.uint64)
        o += ["Smallest key range : %d" % int(d.min()), "Largest key range  : %d" % int(d.max())]
    o += ["", SEP, "Array samples", SEP]
    for name, v in (("P", P), ("Perm", Perm), ("K", K), ("N", N), ("NF", NF)):
        o += sample_lines(name, v) + [""]
    o += ["", SEP, "Sample leaves", SEP]

    def leaf(i):
        if i < len(N) and i + 1 < len(K):
            o.append("leaf[%d]  K=[%d, %d)  range=%d  particles=%d" % (i, int(K[i]), int(K[i + 1]),
                                                                       int(K[i + 1]) - int(K[i]), int(N[i])))
    head = min(4, L)
    for i in range(head):
        leaf(i)
    if L > 8:
        o.append("...")
    if L > 4:
        for i in range(max(head, L - 4), L):
            leaf(i)
    o += ["", SEP]
    box = t.get("sfcBoxIndex_d")
    if len(box):
        o += ["Tree-structure view", SEP, "maxAchievedDepth  : %d" % depth]
        for label, name in (("SFCBoxIndex", "sfcBoxIndex_d"), ("hasBoxSplit", "hasBoxSplit_d"),
                            ("boxDepth", "boxDepth_d"), ("parentIndex", "parentIndex_d"),
                            ("childIndex", "childIndex_d"), ("levelOffset", "levelOffset_d")):
            o.append(array_line(label, t.get(name)))
        o.append("")
    counts = t.get("particleCounts_d")
    if len(counts):
        o += [SEP, "Per-box particle counts", SEP, array_line("particleBeginIdx", t.get("particleBeginIndex_d")),
              array_line("particleCounts", counts), ""]
    mapping = t.get("sfcParticleIndex_d")
    if len(mapping):
        o += [SEP, "Particle <-> leaf mapping", SEP, array_line("SFCParticleIndex", mapping), ""]
    o.append(SEP)
    return "\n".join(o) + "\n"


def binary_text(t, depth):
    K, P, N, perm = t.get("cornerstone_d"), t.get("particleKeys_d"), t.get("leafCounts_d"), t.get("perm_d")
    sc = t.scalars()
    o = ["Adaptive 3D Binary Tree", "LeafCount / longest-side binary keys",
         "Particles: %d / %d input slots" % (sc.num_active_particles, sc.num_particles),
         "Leaves: %d" % len(N), "Nodes: %d" % sc.num_nodes, "Maximum binary depth: %d" % depth]
    if len(N):
        s = int(N.sum(dtype=np.uint64))
        o.append("N: min=%d max=%d sum=%d mean=%s empty=%d" % (N.min(), N.max(), s, "%g" % (s / len(N)),
                                                               int((N == 0).sum())))
    for name, v in (("P", P), ("Perm", perm)):
        items = ", ".join(str(int(x)) for x in v[:min(4, len(v))])
        o.append("%s size=%d samples=[%s%s" % (name, len(v), items, ", ...]" if len(v) > 4 else "]"))
    o.append("K: leaf boundary array (%d keys)" % len(K))

    def leaf(i):
        r = int(K[i + 1]) - int(K[i])
        o.append("  leaf[%d] depth=%d K=[%d,%d) N=%d" % (i, 64 - r.bit_length(), int(K[i]), int(K[i + 1]), int(N[i])))
    head = min(4, len(N))
    for i in range(head):
        leaf(i)
    if len(N) > 8:
        o.append("  ...")
    for i in range(max(head, len(N) - 4 if len(N) > 4 else 0), len(N)):
        leaf(i)
    return "\n".join(o) + "\n"


CASES = [
    ("octree_leafcount_views", OCTREE, dict(max_particles_per_leaf=24, sfc_kind=MORTON, views=(1, 1, 1))),
    ("octree_nf_hilbert", OCTREE, dict(max_particles_per_leaf=4096, split_criterion=NF_COUNT, sfc_kind=HILBERT)),
    ("binary_views", BINARY, dict(max_particles_per_leaf=16, views=(1, 1, 1))),
]


@pytest.mark.parametrize("name,kind,cfg", CASES, ids=[c[0] for c in CASES])
def test_print_matches_the_reference_layout(name, kind, cfg):
    print("print", name, flush=True)
    rng = np.random.default_rng(21)
    n = 30000
    pts = np.concatenate([0.6 + 0.01 * rng.standard_normal((n // 2, 3)), rng.random((n - n // 2, 3))]).clip(0, 1)
    host = [np.ascontiguousarray(pts[:, a]) for a in range(3)]
    for a in range(3):
        host[a][[7, 70]] = np.finfo(np.float64).max
    dev = [DeviceArray(a) for a in host]
    g = Tree(*dev, n, make_config(**cfg), real=np.float64, kind=kind)
    o = OracleTree(*host, n, make_config(**cfg), real=np.float64, kind=kind)
    for t in (g, o):
        t.build()
        if cfg.get("views"):
            t.computeViews()
        t.doBuffersNeedResize()
    text = g.printString()
    if kind == BINARY:
        want = binary_text(o, o.maxAchievedDepth())
    else:
        # a LeafCount octree's NF contents are unset in the reference (its
        # buffer is resized without initialisation); the print shows the
        # getter's values, which only NFCount trees define
        nf = None if cfg.get("split_criterion") == NF_COUNT else g.get("nearFieldCounts_d")
        want = octree_text(o, o.maxAchievedDepth(), nf)
    assert text == want, "\n".join(
        "%d: got %r want %r" % (i, a, b) for i, (a, b) in enumerate(zip(text.splitlines(), want.splitlines()))
        if a != b)
