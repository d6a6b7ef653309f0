"""ctypes bindings of the C ABI declared in include/adaptive_trees.h.

``Lib`` binds the tree entry points of a library by prefix; ``hip_lib()``
loads the HIP engine and raises if it is not built (there is no CPU
fallback).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIP_LIB = os.path.join(HERE, "libadaptive_trees_hip.so")

AT_OK, AT_EINVAL, AT_ELOGIC, AT_ELENGTH, AT_EHIP, AT_ENOMEM, AT_EIO = range(7)
OCTREE, BINARY = 0, 1
LEAF_COUNT, NF_COUNT = 0, 1
MORTON, HILBERT = 0, 1

BUFFERS = {
    "particleKeys_d": 0, "perm_d": 1, "cornerstone_d": 2, "leafCounts_d": 3,
    "nearFieldCounts_d": 4, "sfcBoxIndex_d": 5, "hasBoxSplit_d": 6,
    "boxDepth_d": 7, "parentIndex_d": 8, "childIndex_d": 9, "levelOffset_d": 10,
    "particleBeginIndex_d": 11, "particleCounts_d": 12, "sfcParticleIndex_d": 13,
}
BUFFER_DTYPES = {
    0: "uint64", 1: "uint32", 2: "uint64", 3: "uint32", 4: "uint32", 5: "uint64",
    6: "uint8", 7: "uint8", 8: "int32", 9: "int32", 10: "int32", 11: "uint32",
    12: "uint32", 13: "uint64",
}


class Box(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")]


class Config(ctypes.Structure):
    _fields_ = [
        ("max_depth", ctypes.c_int32),
        ("max_particles_per_leaf", ctypes.c_uint32),
        ("update_iterations", ctypes.c_int32),
        ("split_criterion", ctypes.c_int32),
        ("sfc_kind", ctypes.c_int32),
        ("views", ctypes.c_uint8 * 3),
        ("reserved0", ctypes.c_uint8),
        ("box", Box),
        ("stream", ctypes.c_void_p),
    ]


class Scalars(ctypes.Structure):
    _fields_ = [
        ("num_particles", ctypes.c_uint64),
        ("num_active_particles", ctypes.c_uint64),
        ("num_nodes", ctypes.c_uint64),
        ("num_leaves", ctypes.c_uint64),
        ("leaf_capacity", ctypes.c_uint64),
        ("tree_type", ctypes.c_int32),
        ("split_criterion", ctypes.c_int32),
        ("sfc_kind", ctypes.c_int32),
        ("real_bytes", ctypes.c_int32),
        ("active_buffer", ctypes.c_int32),
        ("update_iterations", ctypes.c_int32),
        ("max_depth", ctypes.c_int32),
        ("max_particles_per_leaf", ctypes.c_uint32),
    ]


class TreeError(RuntimeError):
    """A failing status code; ``code`` is the AT_E* value."""

    def __init__(self, code, message):
        super().__init__("status %d: %s" % (code, message))
        self.code = code


class InvalidArgument(TreeError, ValueError):
    pass


class LogicError(TreeError):
    pass


class LengthError(TreeError):
    pass


class InputError(TreeError):
    """A particle or manifest file that cannot be read (std::runtime_error)."""


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("workers", ctypes.c_int32), ("max_batch", ctypes.c_int32),
                ("update_graphs", ctypes.c_int32), ("reserved0", ctypes.c_int32)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("bytes_in_use", "bytes_peak", "hbm_budget", "resident_trees",
                                                "waiting_builds", "completed", "batched_updates",
                                                "update_batches", "grouped_builds", "grouped_updates")]


_ERRORS = {AT_EINVAL: InvalidArgument, AT_ELOGIC: LogicError, AT_ELENGTH: LengthError, AT_EIO: InputError}

_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_I = ctypes.c_int

TREE_PROTOS = {
    "tree_create": [_I, _I, _P, _P, _P, _SZ, ctypes.POINTER(Config), ctypes.POINTER(_P)],
    "tree_destroy": [_P],
    "tree_build": [_P],
    "tree_update": [_P],
    "tree_compute_views": [_P, _P],
    "tree_needs_resize": [_P, ctypes.POINTER(_I)],
    "tree_resize_buffers": [_P],
    "tree_buffer": [_P, _I, ctypes.POINTER(_P), ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)],
    "tree_scalars": [_P, ctypes.POINTER(Scalars)],
    "tree_device_counters": [_P, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_P)],
    "tree_max_achieved_depth": [_P, ctypes.POINTER(ctypes.c_uint32)],
}



class Lib:
    """A loaded library plus its function prefix."""

    def __init__(self, path, prefix):
        if not os.path.exists(path):
            raise FileNotFoundError(
                "%s is not built; run __graft_entry__.build() first" % path)
        self.path = path
        self.prefix = prefix
        self.dll = ctypes.CDLL(path)
        for name, args in TREE_PROTOS.items():
            fn = getattr(self.dll, prefix + name)
            fn.argtypes = args
            fn.restype = _I
        last = getattr(self.dll, prefix + "last_error")
        last.restype = ctypes.c_char_p
        last.argtypes = []

    def fn(self, name):
        return getattr(self.dll, self.prefix + name)

    def check(self, status):
        if status != AT_OK:
            msg = (getattr(self.dll, self.prefix + "last_error")() or b"").decode(errors="replace")
            raise _ERRORS.get(status, TreeError)(status, msg)

    def call(self, name, *args):
        self.check(self.fn(name)(*args))


_hip = None


def hip_lib():
    """The HIP engine; raises if it has not been built (no CPU fallback)."""
    global _hip
    if _hip is None:
        lib = Lib(HIP_LIB, "at_")
        d = lib.dll
        for name, args in {
            "device_count": [ctypes.POINTER(_I)],
            "device_malloc": [ctypes.POINTER(_P), _SZ],
            "device_free": [_P],
            "memcpy": [_P, _P, _SZ, _I],
            "memcpy_async": [_P, _P, _SZ, _I, _P],
            "stream_create": [ctypes.POINTER(_P)],
            "stream_destroy": [_P],
            "stream_synchronize": [_P],
            "device_synchronize": [],
            "event_create": [ctypes.POINTER(_P)],
            "event_destroy": [_P],
            "event_record": [_P, _P],
            "event_elapsed_ms": [_P, _P, ctypes.POINTER(ctypes.c_float)],
            "graph_capture_update": [_P, _P, ctypes.POINTER(_P)],
            "graph_launch": [_P, _P],
            "graph_destroy": [_P],
            "profile_enable": [_I],
            "profile_reset": [],
            "profile_query": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_ulonglong)],
            "bench_copy": [_P, _P, _SZ, _P],
            "tune_set": [ctypes.c_char_p, _I],
            "tune_get": [ctypes.c_char_p, ctypes.POINTER(_I)],
            "tune_auto": [_SZ, _P, ctypes.c_char_p, _SZ],
            "tree_update_passes": [_P, ctypes.POINTER(ctypes.c_uint32)],
            "sort_pairs": [_P, _P, _SZ, _I, _P],
            "sort_pairs_noalloc": [_P, _P, _SZ, _I, _P],
            "batch_create": [_I, _I, _SZ, _P, _P, _P, _P, ctypes.POINTER(Config), ctypes.POINTER(_P)],
            "batch_destroy": [_P],
            "batch_build": [_P],
            "batch_update": [_P],
            "batch_needs_resize": [_P, _P, ctypes.POINTER(_SZ)],
            "batch_resize_buffers": [_P],
            "batch_size": [_P, ctypes.POINTER(_SZ)],
            "batch_tree_buffer": [_P, _SZ, _I, ctypes.POINTER(_P), ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)],
            "batch_tree_scalars": [_P, _SZ, ctypes.POINTER(Scalars)],
            "batch_graph_capture_update": [_P, _P, ctypes.POINTER(_P)],
            "server_create": [ctypes.POINTER(ServerConfig), ctypes.POINTER(_P)],
            "server_destroy": [_P],
            "server_submit_build": [_P, _I, _I, _P, _P, _P, _SZ, ctypes.POINTER(Config),
                                    ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)],
            "server_submit_update": [_P, ctypes.c_uint64, _P, _P, _P, ctypes.POINTER(ctypes.c_uint64)],
            "server_submit_release": [_P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)],
            "server_wait": [_P, ctypes.c_uint64, ctypes.POINTER(_I)],
            "server_poll": [_P, ctypes.c_uint64, ctypes.POINTER(_I)],
            "server_request_times": [_P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
            "server_tree_buffer": [_P, ctypes.c_uint64, _I, ctypes.POINTER(_P), ctypes.POINTER(_SZ),
                                   ctypes.POINTER(_SZ)],
            "server_tree_scalars": [_P, ctypes.c_uint64, ctypes.POINTER(Scalars)],
            "server_get_stats": [_P, ctypes.POINTER(ServerStats)],
            "tree_print": [_P, ctypes.c_char_p, _SZ, ctypes.POINTER(_SZ)],
            "tree_leaf_count_violations": [_P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64),
                                           ctypes.POINTER(ctypes.c_uint64)],
            "io_read_columns": [ctypes.c_char_p, _I, ctypes.POINTER(ctypes.POINTER(ctypes.c_double)),
                                ctypes.POINTER(_SZ)],
            "io_read_frame": [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(ctypes.c_double)),
                              ctypes.POINTER(_SZ)],
            "io_manifest_open": [ctypes.c_char_p, ctypes.POINTER(_P)],
            "io_manifest_info": [_P, ctypes.c_char_p, _SZ, ctypes.POINTER(_SZ), ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(_SZ)],
            "io_manifest_frame": [_P, _SZ, ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_long), ctypes.c_char_p, _SZ,
                                  ctypes.c_char_p, _SZ],
        }.items():
            f = getattr(d, "at_" + name)
            f.argtypes = args
            f.restype = _I
        for name, args in {"io_free": [_P], "io_manifest_close": [_P]}.items():
            f = getattr(d, "at_" + name)
            f.argtypes = args
            f.restype = None
        _hip = lib
    return _hip
