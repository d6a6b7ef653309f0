"""Particle input formats (host only, no GPU): the test reader of
tests/particle_io.hpp:35-139 and the harness manifest/frames of
benchmarks/drivers/common.hpp:54-175, through the C ABI (io.cpp).

Pinned against the reference's own data: the head of its verification
snapshot (tests/golden/snapshot_0000_head.dat, the reference's bytes) must
parse to the coordinates in reference_fixture.npz, which were parsed with
Python's correctly rounded float() -- the value strtod and from_chars give.
The edge cases follow the readers' rules line by line: blank and '#' rows,
ESC rows (and what may follow them), separators after numbers, the two
column layouts, binary frames with NaN = escaped, and manifest paths.
"""
import os

import numpy as np
import pytest

from adaptive_trees_amd import _ffi, io

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
E = io.ESCAPED


@pytest.fixture(scope="module")
def fixture():
    return np.load(os.path.join(GOLDEN, "reference_fixture.npz"))


def write(tmp_path, name, text, mode="w"):
    p = tmp_path / name
    with open(p, mode) as f:
        f.write(text)
    return str(p)


def test_reference_snapshot_head(fixture):
    path = os.path.join(GOLDEN, "snapshot_0000_head.dat")
    x, y, z = io.read_particles(path, io.CHARGE_XYZ)
    n = len(x)
    assert n == 1500
    assert np.array_equal(x, fixture["x0000"][:n])
    assert np.array_equal(y, fixture["y0000"][:n])
    assert np.array_equal(z, fixture["z0000"][:n])
    # the harness reader of the same rows as a "cxyz" frame
    fx, fy, fz = io.read_frame(path, "cxyz")
    assert np.array_equal(fx, x) and np.array_equal(fy, y) and np.array_equal(fz, z)


def test_columns_skip_comments_blank_and_escape(tmp_path):
    p = write(tmp_path, "a.dat", "# header\n\n   \n1 2 3 4\n\tESC\t# gone\n  -5e-1 +6 0x1p3 inf\n"
                                 "7 8 9 10 11 extra\n")
    c = io.read_columns(p, 4)
    assert c.shape == (4, 4)
    assert list(c[:, 0]) == [1, 2, 3, 4]
    assert list(c[:, 1]) == [E, E, E, E]
    # strtod: leading '+', hexadecimal and "inf" are numbers
    assert c[0, 2] == -0.5 and c[1, 2] == 6 and c[2, 2] == 8 and np.isinf(c[3, 2])
    # columns past the requested count are not read
    assert list(c[:, 3]) == [7, 8, 9, 10]
    # a number may end at '#', which ends the row's numbers
    q = write(tmp_path, "b.dat", "1e2# comment\t\n3 4\n")
    one = io.read_columns(q, 1)
    assert one.shape == (1, 2) and list(one[0]) == [100, 3]
    with pytest.raises(_ffi.InputError):
        io.read_columns(q, 2)


@pytest.mark.parametrize("row", ["1 2 x 4", "1,2,3,4", "1 2 3", "ESC junk", "ESCAPE", "1 2 3 4x", "--1 2 3 4"])
def test_invalid_rows_name_the_line(tmp_path, row):
    p = write(tmp_path, "bad.dat", "1 2 3 4\n# c\n" + row + "\n")
    with pytest.raises(_ffi.InputError) as e:
        io.read_columns(p, 4)
    assert e.value.code == _ffi.AT_EIO
    assert "bad.dat:3" in str(e.value)


def test_column_count_and_missing_file(tmp_path):
    p = write(tmp_path, "a.dat", "1\n")
    for bad in (0, 17):
        with pytest.raises(_ffi.InvalidArgument):
            io.read_columns(p, bad)
    with pytest.raises(_ffi.InputError):
        io.read_columns(str(tmp_path / "missing.dat"), 4)
    with pytest.raises(_ffi.InvalidArgument):
        io.read_particles(p, "zyx")


def test_layouts(tmp_path):
    p = write(tmp_path, "l.dat", "9 1 2 3\nESC\n8 4 5 6\n")
    x, y, z = io.read_particles(p, io.CHARGE_XYZ)
    assert list(x) == [1, E, 4] and list(y) == [2, E, 5] and list(z) == [3, E, 6]
    x, y, z = io.read_particles(p, io.XYZ_CHARGE)
    assert list(x) == [9, E, 8] and list(y) == [1, E, 4] and list(z) == [2, E, 5]


def test_text_frames(tmp_path):
    # the harness reader: one leading '+', "ESC" followed by anything, extra
    # columns ignored, numbers need not be followed by a separator check
    p = write(tmp_path, "f.dat", "# c\n1 +2 3 4 tail\nESC whatever 1 2\n  5\t6 7 8\n\n")
    x, y, z = io.read_frame(p, "cxyz")
    assert list(x) == [2, E, 6] and list(y) == [3, E, 7] and list(z) == [4, E, 8]
    x, y, z = io.read_frame(p, "xyzc")
    assert list(x) == [1, E, 5] and list(y) == [2, E, 6] and list(z) == [3, E, 7]
    bad = write(tmp_path, "g.dat", "1 2 3 4\n1 2 three 4\n")
    with pytest.raises(_ffi.InputError) as e:
        io.read_frame(bad, "cxyz")
    assert "g.dat:2" in str(e.value)
    with pytest.raises(_ffi.InputError):
        io.read_frame(p, "csv")


def test_binary_frames(tmp_path):
    v = np.array([[1.0, 2.0, 3.0], [np.nan, 5.0, 6.0], [7.0, np.nan, 9.0], [-1e300, 0.0, 2.5]])
    p = str(tmp_path / "f.bin")
    v.astype(np.float64).tofile(p)
    x, y, z = io.read_frame(p, "bin")
    # NaN in x marks an escaped slot; other NaNs are kept
    assert list(x) == [1.0, E, 7.0, -1e300]
    assert list(y[[0, 1, 3]]) == [2.0, E, 0.0] and np.isnan(y[2])
    assert list(z) == [3.0, E, 9.0, 2.5]
    with open(p, "ab") as f:
        f.write(b"\0" * 8)
    with pytest.raises(_ffi.InputError):
        io.read_frame(p, "bin")
    empty = str(tmp_path / "e.bin")
    open(empty, "wb").close()
    assert len(io.read_frame(empty, "bin")[0]) == 0


def test_manifest(tmp_path):
    d = tmp_path / "ds"
    d.mkdir()
    (d / "frames").mkdir()
    np.zeros((3, 3)).tofile(d / "frames" / "noisy_001.bin")
    # the layout written by the reference's snapshot generator
    # (benchmarks/gen_snapshots.py:83-89)
    text = ("dataset halo\nslots 3\nbox -0.5 1.5 -1.0000000000000002 2 0 1\n"
            "frame 0 0 xyzc ../src/halo.dat\nframe 1 7 bin frames/noisy_001.bin\n"
            "frame 2 9 cxyz /abs/path with space.dat\nunknown tag ignored\n")
    m = io.read_manifest(write(d, "manifest.txt", text))
    assert m.dataset == "halo" and m.slots == 3
    assert m.box == (-0.5, 1.5, -1.0000000000000002, 2.0, 0.0, 1.0)
    assert [(f.index, f.id, f.format) for f in m.frames] == [(0, 0, "xyzc"), (1, 7, "bin"), (2, 9, "cxyz")]
    assert m.frames[0].path == str(tmp_path / "src" / "halo.dat")
    assert m.frames[1].path == str(d / "frames" / "noisy_001.bin")
    assert m.frames[2].path == "/abs/path with space.dat"
    x, _, _ = io.read_frame(m.frames[1])
    assert len(x) == 3
    for bad in ("dataset x\nslots 3\n", "dataset x\nslots 0\nframe 0 0 bin a.bin\n"):
        with pytest.raises(_ffi.InputError):
            io.read_manifest(write(d, "bad.txt", bad))
    with pytest.raises(_ffi.InputError):
        io.read_manifest(str(d / "none.txt"))


def test_large_files_parse_in_chunks(tmp_path):
    """Files of many MiB are parsed by several threads in row-aligned chunks:
    the rows come out in file order and a bad row far into the file is
    reported with its line number in the whole file."""
    rng = np.random.default_rng(3)
    n = 260_000
    v = rng.standard_normal((n, 4)) * 10.0 ** rng.integers(-30, 30, (n, 1))
    esc = rng.random(n) < 0.01
    rows = ["ESC" if e else "%.17g %.17g %.17g %.17g" % tuple(r) for r, e in zip(v, esc)]
    rows.insert(1000, "# a comment")
    rows.insert(200_000, "")
    p = write(tmp_path, "big.dat", "\n".join(rows) + "\n")
    assert os.path.getsize(p) > (16 << 20)
    c = io.read_columns(p, 4)
    want = np.where(esc[:, None], E, v).T
    assert c.shape == (4, n) and np.array_equal(c, want)
    x, y, z = io.read_frame(p, "cxyz")
    assert np.array_equal(x, want[1]) and np.array_equal(z, want[3])
    rows[250_000] = "1 2 3 oops"
    q = write(tmp_path, "bad_big.dat", "\n".join(rows) + "\n")
    with pytest.raises(_ffi.InputError) as e:
        io.read_columns(q, 4)
    assert "bad_big.dat:250001" in str(e.value)
    with pytest.raises(_ffi.InputError) as e:
        io.read_frame(q, "xyzc")
    assert "bad_big.dat:250001" in str(e.value)
