"""The branch-free binary key (binaryKeyFast) against the general form
(binaryKey), both from adaptive_trees_amd/csrc/common.hpp, on the host.

The general form follows the axis schedule of makeBinaryGeometry
(binary_node.cuh:17-59) and is checked against the oracle by the GPU parity
tests; the key generation kernel takes the fast form whenever the geometry
allows it. This test compiles tests/cpp/binary_key_forms.hip (host code
only runs) and requires equal keys over random eligible boxes in FP32 and
FP64, including coordinates on and outside the box faces, NaN, infinities
and escaped slots.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not available")
def test_fast_and_general_binary_keys_agree(tmp_path):
    exe = str(tmp_path / "binary_key_forms")
    cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc", "-O1", "-std=c++20", "-ffp-contract=off",
           "--offload-arch=gfx950", "-I", os.path.join(ROOT, "adaptive_trees_amd", "csrc"),
           "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "binary_key_forms.hip"),
           "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    boxes = int(r.stdout.split()[1])
    assert boxes > 500, r.stdout  # most random boxes admit the fast form
