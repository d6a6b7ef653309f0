"""Python mirror of ``adaptive_octree::Tree<Real, TreeType>`` over the C ABI.

Method names, argument meaning and error behaviour follow the reference
facade (include/adaptive_octree/tree.hpp:11-134): build(), update(),
resizeBuffers(), computeViews(), doBuffersNeedResize(), the device counters
and the P/Perm/K/N/NF/view getters. Getters return host copies (numpy) of
the device buffers, trimmed to the logical sizes of the last
build()/doBuffersNeedResize(), exactly as the reference's thrust vectors are.

Invalid configurations raise ``InvalidArgument`` (a ValueError), wrong call
order raises ``LogicError`` and capacity overflow ``LengthError``, the
Python counterparts of the reference's std exceptions.
"""
import ctypes

import numpy as np

from . import _ffi
from ._ffi import BINARY, HILBERT, LEAF_COUNT, MORTON, NF_COUNT, OCTREE  # noqa: F401

DEFAULT_BOX = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)


def make_config(max_depth=-1, max_particles_per_leaf=64, update_iterations=-1,
                split_criterion=LEAF_COUNT, sfc_kind=MORTON, views=(False, False, False),
                box=DEFAULT_BOX, stream=None):
    """TreeConfig<Real> (include/adaptive_octree/config.hpp:143-155)."""
    c = _ffi.Config()
    c.max_depth = max_depth
    c.max_particles_per_leaf = max_particles_per_leaf
    c.update_iterations = update_iterations
    c.split_criterion = split_criterion
    c.sfc_kind = sfc_kind
    for i in range(3):
        c.views[i] = 1 if views[i] else 0
    c.box = _ffi.Box(*[float(v) for v in box])
    c.stream = stream
    return c


class DeviceArray:
    """A device allocation owned from Python (plumbing for tests and bench)."""

    def __init__(self, nbytes_or_array, dtype=None):
        lib = _ffi.hip_lib()
        self._lib = lib
        if isinstance(nbytes_or_array, np.ndarray):
            arr = np.ascontiguousarray(nbytes_or_array)
            self.dtype = arr.dtype
            self.shape = arr.shape
            self.nbytes = arr.nbytes
        else:
            self.dtype = np.dtype(dtype or np.uint8)
            self.nbytes = int(nbytes_or_array)
            self.shape = (self.nbytes // self.dtype.itemsize,)
            arr = None
        p = ctypes.c_void_p()
        lib.call("device_malloc", ctypes.byref(p), max(self.nbytes, 1))
        self.ptr = p.value
        if arr is not None:
            self.upload(arr)

    def upload(self, arr, stream=None):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes <= self.nbytes
        if arr.nbytes:
            self._lib.call("memcpy", self.ptr, arr.ctypes.data, arr.nbytes, 1)

    def download(self, count=None):
        n = self.shape[0] if count is None else count
        out = np.empty(n, dtype=self.dtype)
        if out.nbytes:
            self._lib.call("memcpy", out.ctypes.data, self.ptr, out.nbytes, 2)
        return out

    def free(self):
        if self.ptr:
            self._lib.call("device_free", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return int(a)


class Tree:
    """``Tree<Real, Kind>`` on the HIP engine. Coordinates are device
    pointers (DeviceArray or int addresses) that the tree borrows and
    re-reads on every update()."""

    def __init__(self, x, y, z, count, config=None, real=np.float64, kind=OCTREE):
        self.lib = self._open_lib()
        self.real = np.dtype(real)
        self.kind = kind
        self._coords = (x, y, z)  # keep alive: the tree borrows them
        self.config = config if config is not None else make_config()
        h = ctypes.c_void_p()
        self.lib.call("tree_create", self.real.itemsize, kind, _ptr(x), _ptr(y), _ptr(z),
                      int(count), ctypes.byref(self.config), ctypes.byref(h))
        self.handle = h.value

    def _open_lib(self):
        return _ffi.hip_lib()

    def _copy_out(self, dst, src, nbytes):
        """Device-to-host copy of a tree buffer."""
        self.lib.call("memcpy", dst, src, nbytes, 2)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.fn("tree_destroy")(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- operations (tree.hpp:35-41) ------------------------------------
    @staticmethod
    def type_of(kind):
        return kind

    def build(self):
        self.lib.call("tree_build", self.handle)

    def update(self):
        self.lib.call("tree_update", self.handle)

    def resizeBuffers(self):
        self.lib.call("tree_resize_buffers", self.handle)

    def computeViews(self, views=None):
        if views is None:
            self.lib.call("tree_compute_views", self.handle, None)
        else:
            v = (ctypes.c_uint8 * 3)(*[1 if f else 0 for f in views])
            self.lib.call("tree_compute_views", self.handle, ctypes.cast(v, ctypes.c_void_p))

    def doBuffersNeedResize(self):
        o = ctypes.c_int()
        self.lib.call("tree_needs_resize", self.handle, ctypes.byref(o))
        return bool(o.value)

    # ---- queries (tree.hpp:43-61) -----------------------------------------
    def scalars(self):
        s = _ffi.Scalars()
        self.lib.call("tree_scalars", self.handle, ctypes.byref(s))
        return s

    def numParticles(self):
        return self.scalars().num_particles

    def numActiveParticles(self):
        return self.scalars().num_active_particles

    def numNodes(self):
        return self.scalars().num_nodes

    def leafCapacity(self):
        return self.scalars().leaf_capacity

    def maxAchievedDepth(self):
        d = ctypes.c_uint32()
        self.lib.call("tree_max_achieved_depth", self.handle, ctypes.byref(d))
        return d.value

    def updatePasses(self):
        """Passes of the last update() that changed the tree (HIP trees only;
        synchronises)."""
        d = ctypes.c_uint32()
        self.lib.call("tree_update_passes", self.handle, ctypes.byref(d))
        return d.value

    def print(self, file=None):
        """print(std::ostream&): the reference's diagnostic dump
        (src/octree/octree_print.cu, src/kdtree3d/kdtree3d_print.cu), written
        to ``file`` (default sys.stdout). HIP trees only."""
        import sys
        (file or sys.stdout).write(self.printString())

    def printString(self):
        """The text of print()."""
        need = ctypes.c_size_t()
        self.lib.call("tree_print", self.handle, None, 0, ctypes.byref(need))
        buf = ctypes.create_string_buffer(max(need.value, 1))
        self.lib.call("tree_print", self.handle, buf, len(buf), ctypes.byref(need))
        return buf.value.decode()

    def leafCountViolations(self, bucket=None):
        """(overfull leaves, mergeable sibling groups) of the leaf-count
        convergence check (benchmarks/drivers/cstone_check.cu:17-43), for
        ``bucket`` particles per leaf (default: the tree's
        maxParticlesPerLeaf). HIP trees only."""
        if bucket is None:
            bucket = self.scalars().max_particles_per_leaf
        over, merge = ctypes.c_uint64(), ctypes.c_uint64()
        self.lib.call("tree_leaf_count_violations", self.handle, int(bucket), ctypes.byref(over), ctypes.byref(merge))
        return over.value, merge.value

    def device_counters(self):
        """(numLeaves, numNodes, numActiveParticles) read from the device."""
        pl, pn, pa = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        self.lib.call("tree_device_counters", self.handle, ctypes.byref(pl), ctypes.byref(pn),
                      ctypes.byref(pa))
        vals = []
        for p, dt in ((pl, np.int32), (pn, np.int32), (pa, np.uint32)):
            out = np.zeros(1, dtype=dt)
            self._copy_out(out.ctypes.data, p.value, 4)
            vals.append(int(out[0]))
        return tuple(vals)

    def splitCriterion(self):
        return self.scalars().split_criterion

    def sfcKind(self):
        if self.kind != OCTREE:
            raise _ffi.LogicError(_ffi.AT_ELOGIC, "Tree: SFC selection requires Octree mode")
        return self.scalars().sfc_kind

    # ---- getters (tree.hpp:62-103) -----------------------------------------
    def buffer_info(self, name):
        bid = _ffi.BUFFERS[name]
        d, s, c = ctypes.c_void_p(), ctypes.c_size_t(), ctypes.c_size_t()
        self.lib.call("tree_buffer", self.handle, bid, ctypes.byref(d), ctypes.byref(s),
                      ctypes.byref(c))
        return d.value, s.value, c.value

    def get(self, name, count=None):
        ptr, size, cap = self.buffer_info(name)
        n = size if count is None else count
        dt = np.dtype(_ffi.BUFFER_DTYPES[_ffi.BUFFERS[name]])
        if not ptr:
            # a size without storage (NF of a LeafCount octree, whose contents
            # the reference leaves unset): zeros
            return np.zeros(n, dtype=dt)
        out = np.empty(n, dtype=dt)
        if n:
            self._copy_out(out.ctypes.data, ptr, out.nbytes)
        return out

    def __getattr__(self, name):
        if name in _ffi.BUFFERS:
            return lambda: self.get(name)
        raise AttributeError(name)
