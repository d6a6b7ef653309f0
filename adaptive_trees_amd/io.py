"""Particle input formats of the reference's tests and benchmark harness.

``read_columns`` / ``read_particles``: the test reader
(tests/particle_io.hpp:35-139) -- whitespace-separated numbers parsed as
strtod does, '#' and blank rows skipped, "ESC" rows kept as escaped slots.

``read_manifest`` / ``read_frame``: the benchmark harness's dataset manifest
and frames (benchmarks/drivers/common.hpp:54-175) -- text frames ("cxyz" =
charge x y z, "xyzc" = x y z charge) and binary frames ("bin" = x y z
doubles per slot, NaN = escaped).

Parsing runs in the native library (io.cpp); results are host numpy
arrays. Escaped slots hold ``ESCAPED`` (the largest double) in every
coordinate, the marker the trees recognise.
"""
import collections
import ctypes
import os

import numpy as np

from . import _ffi

ESCAPED = np.finfo(np.float64).max
CHARGE_XYZ, XYZ_CHARGE = "cxyz", "xyzc"

Frame = collections.namedtuple("Frame", "index id format path")
Manifest = collections.namedtuple("Manifest", "dataset slots box frames")


def _lib():
    return _ffi.hip_lib()


def _take(lib, ptr, count):
    try:
        return np.ctypeslib.as_array(ptr, shape=(count,)).copy() if count else np.zeros(0)
    finally:
        lib.dll.at_io_free(ctypes.cast(ptr, ctypes.c_void_p))


def read_columns(path, columns):
    """``columns`` numbers per row (particle_io::readColumns); returns an
    array [columns, rows] of float64."""
    lib = _lib()
    data = ctypes.POINTER(ctypes.c_double)()
    rows = ctypes.c_size_t()
    lib.call("io_read_columns", os.fsencode(path), int(columns), ctypes.byref(data), ctypes.byref(rows))
    return _take(lib, data, int(columns) * rows.value).reshape(int(columns), rows.value)


def read_particles(path, layout=CHARGE_XYZ):
    """x, y, z of a four-column particle file (particle_io::readParticles):
    layout "cxyz" (charge x y z) or "xyzc" (x y z charge)."""
    if layout not in (CHARGE_XYZ, XYZ_CHARGE):
        raise _ffi.InvalidArgument(_ffi.AT_EINVAL, "particle_io::readParticles: unknown layout")
    c = read_columns(path, 4)
    k = 1 if layout == CHARGE_XYZ else 0
    return c[k].copy(), c[k + 1].copy(), c[k + 2].copy()


def read_frame(frame, format=None):
    """x, y, z of one harness frame (bench::readFrame); ``frame`` is a
    ``Frame`` or a path (then ``format`` is "cxyz", "xyzc" or "bin")."""
    if isinstance(frame, Frame):
        path, format = frame.path, frame.format
    else:
        path = frame
    lib = _lib()
    data = ctypes.POINTER(ctypes.c_double)()
    n = ctypes.c_size_t()
    lib.call("io_read_frame", os.fsencode(path), str(format).encode(), ctypes.byref(data), ctypes.byref(n))
    xyz = _take(lib, data, 3 * n.value).reshape(3, n.value)
    return xyz[0].copy(), xyz[1].copy(), xyz[2].copy()


def read_manifest(path):
    """A dataset manifest (bench::readManifest): dataset name, slot count,
    box (xmin, xmax, ymin, ymax, zmin, zmax) and frames in file order, with
    frame paths resolved against the manifest's directory."""
    lib = _lib()
    h = ctypes.c_void_p()
    lib.call("io_manifest_open", os.fsencode(path), ctypes.byref(h))
    try:
        name = ctypes.create_string_buffer(4096)
        slots, count = ctypes.c_size_t(), ctypes.c_size_t()
        box = (ctypes.c_double * 6)()
        lib.call("io_manifest_info", h, name, len(name), ctypes.byref(slots), box, ctypes.byref(count))
        frames = []
        for i in range(count.value):
            index, fid = ctypes.c_int(), ctypes.c_long()
            fmt, fpath = ctypes.create_string_buffer(64), ctypes.create_string_buffer(8192)
            lib.call("io_manifest_frame", h, i, ctypes.byref(index), ctypes.byref(fid), fmt, len(fmt), fpath,
                     len(fpath))
            frames.append(Frame(index.value, fid.value, fmt.value.decode(), os.fsdecode(fpath.value)))
        return Manifest(name.value.decode(), slots.value, tuple(box), frames)
    finally:
        lib.dll.at_io_manifest_close(h)
