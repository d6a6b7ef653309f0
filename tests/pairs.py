"""Run the HIP engine and the CPU oracle side by side on the same inputs.

Every comparison is exact (integer arrays, bit for bit).
"""
import numpy as np

from adaptive_trees_amd import BINARY, OCTREE, DeviceArray, Tree, make_config
from oracle.binding import OracleTree

ARRAYS = ("particleKeys_d", "perm_d", "cornerstone_d", "leafCounts_d", "nearFieldCounts_d",
          "sfcBoxIndex_d", "hasBoxSplit_d", "boxDepth_d", "parentIndex_d", "childIndex_d",
          "levelOffset_d", "particleBeginIndex_d", "particleCounts_d", "sfcParticleIndex_d")


class Pair:
    """A GPU tree and an oracle tree over the same coordinates."""

    def __init__(self, x, y, z, kind=OCTREE, real=np.float64, **cfg):
        self.real = np.dtype(real)
        self.kind = kind
        x, y, z = (np.ascontiguousarray(a, dtype=self.real) for a in (x, y, z))
        self.n = len(x)
        self.host = [x.copy(), y.copy(), z.copy()]
        self.dev = [DeviceArray(a) for a in self.host]
        self.gpu = Tree(*self.dev, self.n, make_config(**cfg), real=self.real, kind=kind)
        self.cpu = OracleTree(*self.host, self.n, make_config(**cfg), real=self.real, kind=kind)

    def set_coords(self, x, y, z):
        for h, d, a in zip(self.host, self.dev, (x, y, z)):
            h[:] = np.asarray(a, dtype=self.real)
            d.upload(h)

    def build(self):
        self.gpu.build()
        self.cpu.build()

    def update(self, resize_loop=True):
        self.gpu.update()
        self.cpu.update()
        rg, rc = self.gpu.doBuffersNeedResize(), self.cpu.doBuffersNeedResize()
        assert rg == rc, "needsResize differs: gpu %s oracle %s" % (rg, rc)
        attempts = 0
        while resize_loop and rg:
            attempts += 1
            assert attempts < 64
            self.gpu.resizeBuffers()
            self.cpu.resizeBuffers()
            self.gpu.update()
            self.cpu.update()
            rg, rc = self.gpu.doBuffersNeedResize(), self.cpu.doBuffersNeedResize()
            assert rg == rc
        return attempts

    def compare(self, what=""):
        sg, sc = self.gpu.scalars(), self.cpu.scalars()
        for f in ("num_particles", "num_active_particles", "num_nodes", "num_leaves", "leaf_capacity"):
            assert getattr(sg, f) == getattr(sc, f), "%s %s: gpu %d oracle %d" % (
                what, f, getattr(sg, f), getattr(sc, f))
        for name in ARRAYS:
            if name == "nearFieldCounts_d" and (self.kind == BINARY or sg.split_criterion == 0):
                continue
            a, b = self.gpu.get(name), self.cpu.get(name)
            assert a.shape == b.shape, "%s %s: shape gpu %s oracle %s" % (what, name, a.shape, b.shape)
            if not np.array_equal(a, b):
                idx = np.nonzero(a != b)[0]
                raise AssertionError("%s %s: %d mismatches, first at %s: gpu %s oracle %s" % (
                    what, name, len(idx), idx[:4], a[idx[:4]], b[idx[:4]]))
        assert self.gpu.maxAchievedDepth() == self.cpu.maxAchievedDepth(), what
        assert self.gpu.device_counters() == self.cpu.device_counters(), what


def uniform(n, seed, box=(0, 1, 0, 1, 0, 1)):
    rng = np.random.default_rng(seed)
    lo = np.array(box[0::2], dtype=np.float64)
    hi = np.array(box[1::2], dtype=np.float64)
    pts = lo + rng.random((3, n)).T * (hi - lo)
    return pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy()


def plummer(n, seed, a=0.05, centre=0.5):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < 1 or sum(len(o) for o in out) < n:
        u = rng.random(n)
        r = a / np.sqrt(np.maximum(u, 1e-300) ** (-2.0 / 3.0) - 1.0)
        r = r[r <= 20 * a]
        m = len(r)
        cz = 2 * rng.random(m) - 1
        ph = 2 * np.pi * rng.random(m)
        s = np.sqrt(1 - cz * cz)
        out.append(np.stack([r * s * np.cos(ph), r * s * np.sin(ph), r * cz], axis=1))
    p = np.concatenate(out)[:n] + centre
    p = np.clip(p, 0.0, 1.0)
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()


def jitter(coords, sigma, seed, box=(0, 1, 0, 1, 0, 1)):
    rng = np.random.default_rng(seed)
    out = []
    for a, (lo, hi) in zip(coords, ((box[0], box[1]), (box[2], box[3]), (box[4], box[5]))):
        v = a + rng.normal(0, sigma * (hi - lo), size=a.shape)
        v = np.where(v < lo, 2 * lo - v, v)
        v = np.where(v > hi, 2 * hi - v, v)
        out.append(np.clip(v, lo, hi))
    return out
