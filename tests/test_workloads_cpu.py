"""The synthetic-workload generators (at_workload_*, host code) against the
published generator they restate.

splitmix64 and xoshiro256** are restated here in plain Python, pinned by
their known outputs (splitmix64 from state 0; xoshiro256** from the state
{1, 2, 3, 4}, whose first two outputs follow by hand: rotl(2 * 5, 7) * 9 =
11520, then 0 because s[1] becomes 0). The library's raw stream must equal
the restatement for several seeds, its uniforms must be (u >> 11) * 2^-53 of
that stream, and each particle set must have the shape its definition gives.
"""
import numpy as np

from adaptive_trees_amd import workloads

M = (1 << 64) - 1


def splitmix(state):
    state = (state + 0x9E3779B97F4A7C15) & M
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return state, z ^ (z >> 31)


def rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & M


def xoshiro_next(s):
    r = (rotl((s[1] * 5) & M, 7) * 9) & M
    t = (s[1] << 17) & M
    s[2] ^= s[0]
    s[3] ^= s[1]
    s[1] ^= s[2]
    s[0] ^= s[3]
    s[2] ^= t
    s[3] = rotl(s[3], 45)
    return r


def seeded(seed):
    st, out = seed, []
    for _ in range(4):
        st, v = splitmix(st)
        out.append(v)
    return out


def test_restatement_known_outputs():
    st, vals = 0, []
    for _ in range(3):
        st, v = splitmix(st)
        vals.append(v)
    assert vals == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    s = [1, 2, 3, 4]
    assert [xoshiro_next(s) for _ in range(4)] == [11520, 0, 1509978240, 1215971899390074240]


def test_library_stream_equals_restatement():
    from adaptive_trees_amd import _ffi
    import ctypes
    lib = _ffi.hip_lib()
    for seed in (0, 1, 5, 123456789, 2**63 + 11):
        out = np.empty(64, dtype=np.uint64)
        lib.check(lib.fn("workload_bits")(ctypes.c_uint64(seed), ctypes.c_size_t(64),
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        s = seeded(seed)
        assert [int(v) for v in out] == [xoshiro_next(s) for _ in range(64)], seed


def test_uniform_is_top_53_bits():
    n = 1000
    x, y, z = workloads.uniform_cube(n, seed=9, dtype=np.float64)
    s = seeded(9)
    ref = np.array([(xoshiro_next(s) >> 11) * 2.0 ** -53 for _ in range(3 * n)]).reshape(n, 3)
    assert np.array_equal(np.stack([x, y, z], axis=1), ref)  # per particle x, y, z
    lo = (0.0, 4.0, 0.0, 1.0, 0.0, 1.0)
    bx, by, bz = workloads.uniform_cube(n, seed=9, box=lo)
    assert np.array_equal(bx, 4.0 * ref[:, 0]) and np.array_equal(by, ref[:, 1])


def test_plummer_shape():
    x, y, z = (a.astype(np.float64) for a in workloads.plummer(200_000, seed=2, dtype=np.float64))
    for a in (x, y, z):
        assert a.min() >= 0.0 and a.max() <= 1.0
    r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2)
    # enclosed mass of a Plummer sphere: M(r) = r^3 / (r^2 + a^2)^(3/2), truncated at 20a
    a = 0.05
    mass = lambda q: q ** 3 / (q * q + a * a) ** 1.5
    for q in (0.5 * a, a, 2 * a, 5 * a):
        expect = mass(q) / mass(20 * a)
        assert abs(np.mean(r <= q) - expect) < 0.01, q
    assert r.max() <= 20 * a * np.sqrt(3) + 1e-12


def test_two_clusters_shape():
    n = 100_000
    x, y, z = workloads.two_clusters(n, seed=3, dtype=np.float64)
    assert x.min() >= 0 and x.max() <= 4 and y.min() >= 0 and y.max() <= 1 and z.min() >= 0 and z.max() <= 1
    n1 = int(0.4 * n)
    for c, sl in ((1.0, slice(0, n1)), (3.0, slice(n1, 2 * n1))):
        assert abs(x[sl].mean() - c) < 0.005 and abs(np.std(x[sl]) - 0.08) < 0.003
        assert abs(y[sl].mean() - 0.5) < 0.005 and abs(np.std(z[sl]) - 0.08) < 0.003
    rest = x[2 * n1:]
    assert abs(rest.mean() - 2.0) < 0.05 and rest.min() < 0.05 and rest.max() > 3.95


def test_reflect_step_and_sizes():
    n = 50_000
    base = workloads.uniform_cube(n, seed=4)
    moved = workloads.reflect_step(base, 1e-2, seed=44)
    d = moved[0] - base[0]
    assert abs(np.std(d) - 1e-2) < 5e-4 and all(a.min() >= 0 and a.max() <= 1 for a in moved)
    edge = (np.zeros(8), np.ones(8), np.full(8, 0.5))
    e = workloads.reflect_step(edge, 0.3, seed=45)
    assert all(a.min() >= 0 and a.max() <= 1 for a in e)
    sizes = workloads.batch_sizes(4096, seed=4)
    assert sizes.min() >= 1000 and sizes.max() <= 2_000_000
    assert abs(np.log(sizes).mean() - 0.5 * (np.log(1e3) + np.log(2e6))) < 0.1
    kinds = workloads.batch_kinds(4096, seed=4)
    assert set(np.unique(kinds)) == {0, 1} and abs(kinds.mean() - 0.5) < 0.05
