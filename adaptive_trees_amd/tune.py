"""Radix-sort tuning: read, set, or let the native autotuner pick.

Every setting gives bit-identical trees; only the sort's schedule changes
(tile shape, key ranges per pass, histogram source, MSD levels or LSD
passes for large sorts, digit width of the LSD passes of smaller sorts).
"""
import ctypes
import json

from . import _ffi

KEYS = ("sort_threads", "sort_items", "sort_ranges", "sort_digit_array", "sort_prefetch", "sort_msd",
        "sort_digit_bits", "sort_msd_min_keys")


def get():
    lib = _ffi.hip_lib()
    out = {}
    for k in KEYS:
        v = ctypes.c_int()
        lib.call("tune_get", k.encode(), ctypes.byref(v))
        out[k] = v.value
    return out


def set(**values):  # noqa: A001 - mirrors at_tune_set
    lib = _ffi.hip_lib()
    for k, v in values.items():
        lib.call("tune_set", k.encode(), int(v))


def autotune(n, stream=None):
    """Times the candidates on a sort of n random keys; keeps and returns the fastest."""
    buf = ctypes.create_string_buffer(512)
    _ffi.hip_lib().call("tune_auto", int(n), stream, buf, len(buf))
    return json.loads(buf.value.decode())
