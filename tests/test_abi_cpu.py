"""The C-ABI library and its Python host layer, without a GPU.

Checks that libadaptive_trees_hip.so loads and exports every function that
include/adaptive_trees.h declares (no compute calls), that the product
package never reaches the oracle, and the host-side configuration mirror.
"""
import ctypes
import os
import re
import subprocess

import pytest

from adaptive_trees_amd import _ffi, make_config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "adaptive_trees.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(at_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_tree_api():
    names = declared_functions()
    for n in ("at_tree_create", "at_tree_destroy", "at_tree_build", "at_tree_update",
              "at_tree_compute_views", "at_tree_needs_resize", "at_tree_resize_buffers",
              "at_tree_buffer", "at_tree_scalars", "at_tree_device_counters", "at_last_error"):
        assert n in names, n


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_ffi.HIP_LIB), "libadaptive_trees_hip.so is not built"
    dll = ctypes.CDLL(_ffi.HIP_LIB)
    missing = [n for n in declared_functions() if not hasattr(dll, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _ffi.HIP_LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(declared_functions()) <= exported


def test_binding_prototypes_match_the_header():
    """Every entry point the Python layer binds is declared in the header."""
    names = set(declared_functions())
    for n in _ffi.TREE_PROTOS:
        assert "at_" + n in names, n
    assert "at_last_error" in names


def test_last_error_without_gpu_work():
    lib = _ffi.hip_lib()
    msg = lib.dll.at_last_error()
    assert msg is None or isinstance(msg, bytes)


def test_null_arguments_are_rejected():
    """Argument checks run before any device work: a null tree or output pointer
    returns AT_EINVAL (status 1), never a crash."""
    lib = _ffi.hip_lib()
    assert lib.fn("tree_build")(None) == _ffi.AT_EINVAL
    assert lib.fn("tree_update")(None) == _ffi.AT_EINVAL
    assert lib.fn("tree_scalars")(None, None) == _ffi.AT_EINVAL
    with pytest.raises(_ffi.InvalidArgument):
        lib.call("tree_build", None)


def test_product_package_never_names_the_oracle():
    pkg = os.path.join(ROOT, "adaptive_trees_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "liboracle" not in text and "orc_" not in text and "oracle.binding" not in text, f


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(FileNotFoundError):
        _ffi.Lib(str(tmp_path / "libadaptive_trees_hip.so"), "at_")


def test_config_layout_mirrors_the_header():
    """at_config: 5 x 32-bit fields, 3 view flags + 1 pad byte, 6 doubles, the stream pointer."""
    assert ctypes.sizeof(_ffi.Config) == 5 * 4 + 4 + 6 * 8 + 8
    assert _ffi.Config.box.offset == 24
    assert ctypes.sizeof(_ffi.Scalars) == 5 * 8 + 8 * 4
    c = make_config(max_particles_per_leaf=7, views=(1, 0, 1), box=(0, 4, 0, 1, 0, 1))
    assert c.max_particles_per_leaf == 7 and list(c.views) == [1, 0, 1] and c.box.xmax == 4.0
