"""Seeded synthetic particle sets for the benchmark and test workloads.

The generators are the library's at_workload_* (host code): one xoshiro256**
stream per call, seeded through splitmix64(seed); uniform doubles
(u >> 11) * 2^-53, Gaussians by Box-Muller in double; coordinates are made in
float64 and cast to the requested precision at the end (round to nearest).
The same host arrays feed the GPU trees and the oracle.

  uniform_cube   uniform in a box (default [0,1]^3)
  plummer        Plummer sphere, scale radius a, truncated at 20a,
                 centred in the unit box and clipped to it
  two_clusters   4:1:1 box, 40%/40% isotropic Gaussians at x = 1 and 3
                 (sigma = 0.02 of the long side), 20% uniform, clipped
  batch_sizes    log-uniform tree sizes for batches
  batch_kinds    one bit per tree (uniform or Plummer)
  reflect_step   Gaussian displacement with reflecting walls
"""
import ctypes

import numpy as np

from . import _ffi

UNIT_BOX = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
ANISO_BOX = (0.0, 4.0, 0.0, 1.0, 0.0, 1.0)

_DP = ctypes.POINTER(ctypes.c_double)


def _lib():
    return _ffi.hip_lib()


def _box(box):
    return (ctypes.c_double * 6)(*[float(v) for v in box])


def _split(xyz, n, dtype):
    return tuple(np.ascontiguousarray(xyz[a * n:(a + 1) * n], dtype=dtype) for a in range(3))


def _call(name, *args):
    lib = _lib()
    f = lib.fn(name)
    lib.check(f(*args))


def uniform_cube(n, seed=1, dtype=np.float64, box=UNIT_BOX):
    out = np.empty(3 * n, dtype=np.float64)
    _call("workload_uniform", ctypes.c_uint64(seed), ctypes.c_size_t(n), _box(box), out.ctypes.data_as(_DP))
    return _split(out, n, dtype)


def plummer(n, seed=2, a=0.05, dtype=np.float32):
    out = np.empty(3 * n, dtype=np.float64)
    _call("workload_plummer", ctypes.c_uint64(seed), ctypes.c_size_t(n), ctypes.c_double(a), out.ctypes.data_as(_DP))
    return _split(out, n, dtype)


def two_clusters(n, seed=3, dtype=np.float32, box=ANISO_BOX):
    out = np.empty(3 * n, dtype=np.float64)
    _call("workload_two_clusters", ctypes.c_uint64(seed), ctypes.c_size_t(n), _box(box), ctypes.c_double(0.02),
          out.ctypes.data_as(_DP))
    return _split(out, n, dtype)


def batch_sizes(count, seed=4, lo=1000, hi=2_000_000):
    out = np.empty(count, dtype=np.int64)
    _call("workload_log_uniform_sizes", ctypes.c_uint64(seed), ctypes.c_size_t(count), ctypes.c_double(lo),
          ctypes.c_double(hi), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return out


def batch_kinds(count, seed=4):
    """One bit per tree (the stream's lowest bit): 1 = Plummer, 0 = uniform."""
    out = np.empty(count, dtype=np.uint64)
    _call("workload_bits", ctypes.c_uint64(seed), ctypes.c_size_t(count),
          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return (out & np.uint64(1)).astype(np.int64)


def reflect_step(coords, sigma, seed, box=UNIT_BOX, dtype=None):
    """x += N(0, (sigma * side)^2) per axis, reflected at the walls, clamped."""
    n = len(coords[0])
    src = np.concatenate([np.asarray(c, dtype=np.float64) for c in coords]) if n else np.zeros(0)
    out = np.empty(3 * n, dtype=np.float64)
    _call("workload_reflect_step", ctypes.c_uint64(seed), ctypes.c_size_t(n), _box(box), ctypes.c_double(sigma),
          src.ctypes.data_as(_DP), out.ctypes.data_as(_DP))
    return _split(out, n, dtype or np.asarray(coords[0]).dtype)
