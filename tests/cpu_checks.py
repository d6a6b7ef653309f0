"""Independent CPU checks for the oracle: closed forms and brute force.

Nothing here calls the oracle or the HIP engine. Each function restates a
property of the reference algorithm from its definition, in a form unlike
the oracle's (recursion instead of Jacobi passes, breadth-first traversal
instead of sorted layout keys, bit loops instead of spread masks), so a
mistake in the oracle shows up as a mismatch.
"""
from fractions import Fraction

import numpy as np

U64 = np.uint64
ROOT_END = 1 << 63
INVALID = (1 << 64) - 1


def interleave(x, y, z, bits=21):
    """Morton key by a per-bit loop: key = sum over b of x_b 2^(3b+2) + y_b 2^(3b+1) + z_b 2^(3b)
    (reference src/common/sfc_keys.cuh:174-215, x most significant)."""
    x, y, z = (np.asarray(a, dtype=U64) for a in (x, y, z))
    k = np.zeros(x.shape, dtype=U64)
    for b in range(bits):
        s = U64(b)
        k |= ((x >> s) & U64(1)) << U64(3 * b + 2)
        k |= ((y >> s) & U64(1)) << U64(3 * b + 1)
        k |= ((z >> s) & U64(1)) << U64(3 * b)
    return k


def unit_grid(v, bits=21):
    """floor(v * 2^bits) clamped to 2^bits - 1: sfc3D for the unit box with xmin = 0, where
    floor(x*m) - 0*m is exact (sfc_keys.cuh:468-496)."""
    g = np.floor(np.asarray(v, dtype=np.float64) * float(1 << bits))
    return np.minimum(g, float((1 << bits) - 1)).astype(U64)


def lower_bound(keys, v):
    return int(np.searchsorted(keys, np.uint64(v), side="left"))


# ---- converged trees as unique fixed points ----

def octree_fixed_point(P, bucket, max_level=21):
    """Leaves of the converged LeafCount octree: a node is internal iff it holds more than
    `bucket` keys of P and its level is below max_level (the unique fixed point of the
    split/merge passes; benchmarks/drivers/cstone_check.cu:17-23). Depth-first recursion."""
    P = np.asarray(P, dtype=U64)
    out = []

    def rec(start, level):
        rng = 1 << (3 * (21 - level))
        cnt = lower_bound(P, start + rng) - lower_bound(P, start)
        if cnt > bucket and level < max_level:
            child = rng >> 3
            for c in range(8):
                rec(start + c * child, level + 1)
        else:
            out.append(start)

    rec(0, 0)
    out.append(ROOT_END)
    return np.array(out, dtype=U64)


def binary_fixed_point(P, bucket, max_depth=63):
    """Leaves of the converged binary tree: internal iff count > bucket and depth < max_depth
    (tests/tree_test.cu:631-640 expectedLeaves, over key intervals)."""
    P = np.asarray(P, dtype=U64)
    out = []
    stack = [(0, ROOT_END, 0)]
    while stack:
        b, e, d = stack.pop()
        cnt = lower_bound(P, e) - lower_bound(P, b)
        if cnt > bucket and d < max_depth:
            h = (e - b) // 2
            stack.append((b + h, e, d + 1))
            stack.append((b, b + h, d + 1))
        else:
            out.append(b)
    out.append(ROOT_END)
    return np.array(out, dtype=U64)


def decode_morton(k, bits=21):
    x = y = z = 0
    for b in range(bits):
        x |= ((k >> (3 * b + 2)) & 1) << b
        y |= ((k >> (3 * b + 1)) & 1) << b
        z |= ((k >> (3 * b)) & 1) << b
    return x, y, z


def near_field_morton(P, start, level):
    """own * sum of the populations of the 27 periodic same-level cells around the node at
    (start, level), Morton keys (octree_rebalance.cuh:579-634), by brute force over cells."""
    P = np.asarray(P, dtype=U64)
    rng = 1 << (3 * (21 - level))
    w = 1 << (21 - level)
    x, y, z = decode_morton(start)
    own = lower_bound(P, start + rng) - lower_bound(P, start)
    total = 0
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                cx = (x + dx * w) % (1 << 21)
                cy = (y + dy * w) % (1 << 21)
                cz = (z + dz * w) % (1 << 21)
                key = int(interleave([cx], [cy], [cz])[0]) & ~(rng - 1)
                total += lower_bound(P, key + rng) - lower_bound(P, key)
    return own * total, own


def octree_nf_fixed_point(P, bucket):
    """Converged NFCount octree (Morton): internal iff NF(node) > bucket; NF grows towards the
    root, so this is the unique fixed point of the passes."""
    P = np.asarray(P, dtype=U64)
    out = []

    def rec(start, level):
        nf, _ = near_field_morton(P, start, level)
        if nf > bucket and level < 21:
            child = 1 << (3 * (21 - level - 1))
            for c in range(8):
                rec(start + c * child, level + 1)
        else:
            out.append(start)

    rec(0, 0)
    out.append(ROOT_END)
    return np.array(out, dtype=U64)


# ---- physical longest-side bisection (tests/tree_test.cu:571-640) ----

def axis_schedule(box, depth=63):
    """Axis halved at each depth: the strictly longest side, ties to the lowest axis
    (binary_node.cuh:17-40), with exact dyadic halving."""
    side = [box[1] - box[0], box[3] - box[2], box[5] - box[4]]
    out = []
    for _ in range(depth):
        a = 0
        for b in (1, 2):
            if side[b] > side[a]:
                a = b
        out.append(a)
        side[a] /= 2
    return out


def physical_key(p, box):
    """Key of a point by descending 63 physical bisections (tree_test.cu:617-629), in exact
    rational arithmetic (the reference uses long double, exact for its boxes)."""
    lo = [Fraction(box[0]), Fraction(box[2]), Fraction(box[4])]
    hi = [Fraction(box[1]), Fraction(box[3]), Fraction(box[5])]
    p = [Fraction(float(v)) for v in p]
    b, e = 0, ROOT_END
    for _ in range(63):
        a = 0
        for c in (1, 2):
            if hi[c] - lo[c] > hi[a] - lo[a]:
                a = c
        mid = (lo[a] + hi[a]) / 2
        h = (e - b) // 2
        if p[a] >= mid:
            lo[a] = mid
            b = b + h
        else:
            hi[a] = mid
            e = b + h
    return b


# ---- breadth-first view oracle (tests/tree_test.cu:660-778, generalised) ----

def bfs_views(K, P, level_bits, max_level):
    """Child-centric BFS node table of a cornerstone K: at each level, all child-0 nodes in
    parent order, then child-1, ... Returns the view arrays of the reference TreeView."""
    K = np.asarray(K, dtype=U64)
    P = np.asarray(P, dtype=U64)
    kids = 1 << level_bits
    kl = [int(v) for v in K]
    pos = {v: i for i, v in enumerate(kl)}
    nodes = [[0, ROOT_END, 0, -1, -1]]  # begin, end, depth, parent, firstChild
    offsets = []
    lb, le = 0, 1
    for depth in range(max_level + 1):
        offsets.append(lb)
        parents = []
        for i in range(lb, le):
            b, e = nodes[i][0], nodes[i][1]
            j = pos[b]
            if kl[j + 1] != e:
                nodes[i][4] = le + len(parents)
                parents.append(i)
        for c in range(kids):
            for p in parents:
                b, e = nodes[p][0], nodes[p][1]
                w = (e - b) // kids
                nodes.append([b + c * w, b + (c + 1) * w, depth + 1, p, -1])
        lb, le = le, len(nodes)
    offsets.append(len(nodes))
    active = lower_bound(P, ROOT_END)
    n = len(nodes)
    box = np.zeros(n, dtype=U64)
    split = np.zeros(n, dtype=np.uint8)
    depth = np.zeros(n, dtype=np.uint8)
    parent = np.zeros(n, dtype=np.int32)
    child = np.zeros(n, dtype=np.int32)
    begin = np.zeros(n, dtype=np.uint32)
    count = np.zeros(n + 1, dtype=np.uint32)
    mapping = np.full(active, INVALID, dtype=U64)
    for i, (b, e, d, p, fc) in enumerate(nodes):
        sh = level_bits * d
        box[i] = (1 << sh) | (b >> (63 - sh))
        split[i] = fc >= 0
        depth[i] = d
        parent[i] = p
        child[i] = fc
        bb, ee = lower_bound(P, b), lower_bound(P, e)
        begin[i] = bb
        count[i] = ee - bb
        if fc < 0:
            mapping[bb:ee] = i
    count[n] = active
    return {
        "sfcBoxIndex_d": box, "hasBoxSplit_d": split, "boxDepth_d": depth,
        "parentIndex_d": parent, "childIndex_d": child,
        "levelOffset_d": np.array(offsets, dtype=np.int32),
        "particleBeginIndex_d": begin, "particleCounts_d": count,
        "sfcParticleIndex_d": mapping, "maxAchievedDepth": int(max(nd[2] for nd in nodes)),
    }
