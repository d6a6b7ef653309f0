"""Seeded synthetic particle sets for the benchmark and test workloads.

All generators draw from numpy's PCG64 (np.random.default_rng(seed)) in
float64 and cast to the requested precision at the end, so the same seed
gives the same coordinates on every host.

  uniform_cube   uniform in a box (default [0,1]^3)
  plummer        Plummer sphere, scale radius a, truncated at 20a,
                 centred in the unit box and clipped to it
  two_clusters   4:1:1 box, 40%/40% isotropic Gaussians at x = 1 and 3
                 (sigma = 0.02 of the long side), 20% uniform, clipped
  batch_sizes    log-uniform tree sizes for batches
  reflect_step   Gaussian displacement with reflecting walls
"""
import numpy as np

UNIT_BOX = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
ANISO_BOX = (0.0, 4.0, 0.0, 1.0, 0.0, 1.0)


def _split(p, dtype):
    p = np.asarray(p, dtype=np.float64)
    return tuple(np.ascontiguousarray(p[:, a], dtype=dtype) for a in range(3))


def uniform_cube(n, seed=1, dtype=np.float64, box=UNIT_BOX):
    rng = np.random.default_rng(seed)
    lo = np.array(box[0::2])
    hi = np.array(box[1::2])
    return _split(lo + rng.random((n, 3)) * (hi - lo), dtype)


def plummer(n, seed=2, a=0.05, dtype=np.float32):
    rng = np.random.default_rng(seed)
    chunks, have = [], 0
    while have < n:
        m = max(n - have, 1024)
        u = rng.random(m)
        u = np.where(u <= 0, np.finfo(float).tiny, u)
        r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        r = r[r <= 20 * a]
        cz = 2 * rng.random(len(r)) - 1
        ph = 2 * np.pi * rng.random(len(r))
        s = np.sqrt(1 - cz * cz)
        chunks.append(np.stack([r * s * np.cos(ph), r * s * np.sin(ph), r * cz], axis=1))
        have += len(r)
    p = np.clip(np.concatenate(chunks)[:n] + 0.5, 0.0, 1.0)
    return _split(p, dtype)


def two_clusters(n, seed=3, dtype=np.float32, box=ANISO_BOX):
    rng = np.random.default_rng(seed)
    lx = box[1] - box[0]
    sigma = 0.02 * lx
    n1 = int(0.4 * n)
    n2 = int(0.4 * n)
    n3 = n - n1 - n2
    c1 = np.array([box[0] + 0.25 * lx, 0.5 * (box[2] + box[3]), 0.5 * (box[4] + box[5])])
    c2 = np.array([box[0] + 0.75 * lx, 0.5 * (box[2] + box[3]), 0.5 * (box[4] + box[5])])
    lo = np.array(box[0::2])
    hi = np.array(box[1::2])
    p = np.concatenate([c1 + sigma * rng.standard_normal((n1, 3)),
                        c2 + sigma * rng.standard_normal((n2, 3)),
                        lo + rng.random((n3, 3)) * (hi - lo)])
    p = np.clip(p, lo, hi)
    return _split(p, dtype)


def batch_sizes(count, seed=4, lo=1000, hi=2_000_000):
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))).astype(np.int64)


def reflect_step(coords, sigma, seed, box=UNIT_BOX, dtype=None):
    """x += N(0, (sigma * side)^2) per axis, reflected at the walls, clamped."""
    rng = np.random.default_rng(seed)
    out = []
    for a, c in enumerate(coords):
        lo, hi = box[2 * a], box[2 * a + 1]
        v = np.asarray(c, dtype=np.float64) + rng.normal(0.0, sigma * (hi - lo), size=len(c))
        v = np.where(v < lo, 2 * lo - v, v)
        v = np.where(v > hi, 2 * hi - v, v)
        out.append(np.ascontiguousarray(np.clip(v, lo, hi), dtype=dtype or np.asarray(c).dtype))
    return tuple(out)
