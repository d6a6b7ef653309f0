"""adaptive_trees_amd: MI355X-native adaptive octree / binary tree engine.

The engine is libadaptive_trees_hip.so (C ABI in include/adaptive_trees.h,
C++ facade in include/adaptive_octree/tree.hpp). This package holds the
ctypes binding and a Python mirror of the reference facade.
"""
from ._ffi import (BINARY, HILBERT, LEAF_COUNT, MORTON, NF_COUNT, OCTREE,  # noqa: F401
                   InvalidArgument, LengthError, LogicError, TreeError)
from .tree import DeviceArray, Tree, make_config  # noqa: F401
from .batch import Batch  # noqa: F401
from .server import Server  # noqa: F401
