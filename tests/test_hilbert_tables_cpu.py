"""The table-driven Hilbert curve (a 48-state automaton derived at compile
time from the bitwise curve) against the bitwise encode and decode of
adaptive_trees_amd/csrc/common.hpp, on the host (tests/cpp/hilbert_tables.cpp):
equal keys and exact inverses on every octant of every level and on 2M
random 21-bit points. The bitwise form itself is pinned against the oracle
(bijection, unit steps, aligned cubes) by tests/test_oracle_cpu.py and the
GPU parity tests."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not available")
def test_table_and_bitwise_hilbert_agree(tmp_path):
    exe = str(tmp_path / "hilbert_tables")
    src = str(tmp_path / "hilbert_tables.hip")
    shutil.copy(os.path.join(ROOT, "tests", "cpp", "hilbert_tables.cpp"), src)
    cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc", "-O2", "-std=c++20", "--offload-arch=gfx950",
           "-I", os.path.join(ROOT, "adaptive_trees_amd", "csrc"), src, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert int(r.stdout.split()[1]) > 2_000_000 and int(r.stdout.split()[3]) == 0, r.stdout
