"""Batches: many independent octrees built and updated with shared launches.

``Batch`` wraps the at_batch_* entry points. Member k borrows the device
coordinates of DeviceArrays (or raw device addresses) xs[k], ys[k], zs[k];
all members share one configuration. ``Batch.tree(k)`` gives a read-only view
of member k with the getters of ``Tree`` (``get``, ``scalars``, ``numNodes``
...), whose results equal those of a ``Tree`` of the same particles driven
through the same build()/update() calls.
"""
import ctypes

import numpy as np

from . import _ffi
from .tree import OCTREE, _ptr


class BatchMember:
    """Getters of one member tree (reference getters, tree.hpp:62-103)."""

    def __init__(self, batch, index):
        self.batch, self.index = batch, index

    def scalars(self):
        s = _ffi.Scalars()
        self.batch.lib.call("batch_tree_scalars", self.batch.handle, self.index, ctypes.byref(s))
        return s

    def buffer_info(self, name):
        d, s, c = ctypes.c_void_p(), ctypes.c_size_t(), ctypes.c_size_t()
        self.batch.lib.call("batch_tree_buffer", self.batch.handle, self.index, _ffi.BUFFERS[name],
                            ctypes.byref(d), ctypes.byref(s), ctypes.byref(c))
        return d.value, s.value, c.value

    def get(self, name):
        ptr, size, _ = self.buffer_info(name)
        out = np.empty(size, dtype=np.dtype(_ffi.BUFFER_DTYPES[_ffi.BUFFERS[name]]))
        if size and ptr:
            self.batch.lib.call("memcpy", out.ctypes.data, ptr, out.nbytes, 2)
        return out

    def numNodes(self):
        return self.scalars().num_nodes

    def numActiveParticles(self):
        return self.scalars().num_active_particles

    def leafCapacity(self):
        return self.scalars().leaf_capacity


class Batch:
    def __init__(self, xs, ys, zs, counts, config, real=np.float32, kind=OCTREE):
        self.lib = _ffi.hip_lib()
        self.real = np.dtype(real)
        self._coords = (xs, ys, zs)
        self.config = config
        n = len(counts)
        arr = ctypes.c_void_p * max(n, 1)
        self._x = arr(*[_ptr(a) for a in xs])
        self._y = arr(*[_ptr(a) for a in ys])
        self._z = arr(*[_ptr(a) for a in zs])
        self._counts = (ctypes.c_size_t * max(n, 1))(*[int(c) for c in counts])
        h = ctypes.c_void_p()
        self.lib.call("batch_create", self.real.itemsize, kind, n, ctypes.cast(self._x, ctypes.c_void_p),
                      ctypes.cast(self._y, ctypes.c_void_p), ctypes.cast(self._z, ctypes.c_void_p),
                      ctypes.cast(self._counts, ctypes.c_void_p), ctypes.byref(self.config), ctypes.byref(h))
        self.handle = h.value
        self.size = n

    def close(self):
        if getattr(self, "handle", None):
            self.lib.fn("batch_destroy")(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build(self):
        self.lib.call("batch_build", self.handle)

    def update(self):
        self.lib.call("batch_update", self.handle)

    def needsResize(self):
        """Per-member needsResize flags (numpy int array)."""
        flags = np.zeros(max(self.size, 1), dtype=np.int32)
        count = ctypes.c_size_t()
        self.lib.call("batch_needs_resize", self.handle, flags.ctypes.data, ctypes.byref(count))
        return flags[:self.size]

    def resizeBuffers(self):
        self.lib.call("batch_resize_buffers", self.handle)

    def capture_update(self, stream):
        g = ctypes.c_void_p()
        self.lib.call("batch_graph_capture_update", self.handle, stream, ctypes.byref(g))
        return g.value

    def tree(self, k):
        if not 0 <= k < self.size:
            raise IndexError(k)
        return BatchMember(self, k)
