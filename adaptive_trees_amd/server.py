"""A serving loop for many resident trees (at_server_* entry points).

``Server`` keeps trees resident in HBM and runs build / update / release
requests from any number of callers: builds are admitted in arrival order
under an HBM budget, each tree is pinned to one worker stream, and updates
of different trees queued on a worker are launched together. Requests return
ids at once; ``wait`` blocks until one is done and returns its status.
"""
import atexit
import ctypes
import weakref

import numpy as np

from . import _ffi
from .tree import OCTREE, _ptr

# Servers still open when the interpreter exits are closed first: their worker
# threads keep launching kernels on borrowed device arrays, which interpreter
# teardown may free (and the HIP runtime's own exit handlers must not run
# under them). Closing drains every queued request and joins the workers.
_LIVE = weakref.WeakSet()


@atexit.register
def _close_live_servers():
    for s in list(_LIVE):
        try:
            s.close()
        except Exception:
            pass


class ServerTree:
    """Getters of a resident tree (valid once its last request is done)."""

    def __init__(self, server, tree_id):
        self.server, self.id = server, tree_id

    def scalars(self):
        s = _ffi.Scalars()
        self.server.lib.call("server_tree_scalars", self.server.handle, self.id, ctypes.byref(s))
        return s

    def get(self, name):
        d, s, c = ctypes.c_void_p(), ctypes.c_size_t(), ctypes.c_size_t()
        self.server.lib.call("server_tree_buffer", self.server.handle, self.id, _ffi.BUFFERS[name],
                             ctypes.byref(d), ctypes.byref(s), ctypes.byref(c))
        out = np.empty(s.value, dtype=np.dtype(_ffi.BUFFER_DTYPES[_ffi.BUFFERS[name]]))
        if s.value and d.value:
            self.server.lib.call("memcpy", out.ctypes.data, d.value, out.nbytes, 2)
        return out


class Server:
    def __init__(self, hbm_budget_bytes, workers=4, max_batch=16, update_graphs=False):
        """update_graphs: capture each tree's update() at its first update and
        replay it as one launch. While a capture runs, HIP fails operations on
        the legacy default stream in every thread of the process, so enable it
        only where no thread uses that stream while requests are in flight."""
        self.lib = _ffi.hip_lib()
        cfg = _ffi.ServerConfig(int(hbm_budget_bytes), int(workers), int(max_batch), 1 if update_graphs else 0, 0)
        h = ctypes.c_void_p()
        self.lib.call("server_create", ctypes.byref(cfg), ctypes.byref(h))
        self.handle = h.value
        self._keep = []
        _LIVE.add(self)

    def close(self):
        """Finishes the requests given to workers, then destroys the server;
        the device arrays it borrowed may be freed afterwards."""
        if getattr(self, "handle", None):
            self.lib.fn("server_destroy")(self.handle)
            self.handle = None
        _LIVE.discard(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit_build(self, x, y, z, count, config, real=np.float64, kind=OCTREE):
        """Returns (tree_id, request_id). The tree borrows x, y, z (device)."""
        self._keep.append((x, y, z, config))
        tree, req = ctypes.c_uint64(), ctypes.c_uint64()
        self.lib.call("server_submit_build", self.handle, np.dtype(real).itemsize, kind, _ptr(x), _ptr(y), _ptr(z),
                      int(count), ctypes.byref(config), ctypes.byref(tree), ctypes.byref(req))
        return tree.value, req.value

    def submit_update(self, tree, new=None):
        """new = (x, y, z) device arrays copied into the tree's coordinates first."""
        req = ctypes.c_uint64()
        nx, ny, nz = (None, None, None) if new is None else (_ptr(a) for a in new)
        if new is not None:
            self._keep.append(new)
        self.lib.call("server_submit_update", self.handle, tree, nx, ny, nz, ctypes.byref(req))
        return req.value

    def submit_release(self, tree):
        req = ctypes.c_uint64()
        self.lib.call("server_submit_release", self.handle, tree, ctypes.byref(req))
        return req.value

    def wait(self, request):
        st = ctypes.c_int()
        self.lib.call("server_wait", self.handle, request, ctypes.byref(st))
        return st.value

    def poll(self, request):
        d = ctypes.c_int()
        self.lib.call("server_poll", self.handle, request, ctypes.byref(d))
        return bool(d.value)

    def times(self, request):
        """(submit, start, done) in ms since the server started."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self.lib.call("server_request_times", self.handle, request, ctypes.byref(a), ctypes.byref(b),
                      ctypes.byref(c))
        return a.value, b.value, c.value

    def tree(self, tree_id):
        return ServerTree(self, tree_id)

    def stats(self):
        s = _ffi.ServerStats()
        self.lib.call("server_get_stats", self.handle, ctypes.byref(s))
        return {f: getattr(s, f) for f, _ in _ffi.ServerStats._fields_}
