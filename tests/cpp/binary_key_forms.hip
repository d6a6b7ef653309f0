// Host check of the two forms of the binary (longest-side) key: for random
// boxes whose geometry admits the branch-free form (binaryKeyFast), both
// forms must give the same key for coordinates inside the box, on its
// faces, just outside, far outside, NaN, infinite and escaped (the largest
// finite value on all three axes). Prints one summary line; exit status 1
// on any mismatch. Built and run by tests/test_binary_key_forms_cpu.py.
#include <cstdio>
#include <limits>
#include <random>

#include "tree_impl.hpp"

using at::Key;

template <class T>
int check(std::mt19937_64 &rng, long long &boxes, long long &keys)
{
    std::uint32_t spread[256];
    for (unsigned v = 0; v < 256; ++v) spread[v] = std::uint32_t(at::spread3(v));
    std::uniform_real_distribution<double> U(0, 1);
    int bad = 0;
    for (int b = 0; b < 2000; ++b)
    {
        double lo[3], len[3];
        for (int a = 0; a < 3; ++a)
        {
            lo[a] = (U(rng) - 0.5) * std::pow(10.0, int(U(rng) * 6) - 3);
            const int kind = int(U(rng) * 4);
            len[a] = kind == 0 ? 1.0
                                : (kind == 1 ? std::pow(2.0, int(U(rng) * 12) - 6) : std::pow(10.0, (U(rng) - 0.5) * 4));
        }
        const at_box box{lo[0], lo[0] + len[0], lo[1], lo[1] + len[1], lo[2], lo[2] + len[2]};
        at::BinaryGeom g;
        try
        {
            g = at::impl::makeBinaryGeom<T>(box, at::kBinLevels);
        }
        catch (...)
        {
            continue;
        }
        if (!g.fast) continue;
        ++boxes;
        for (int i = 0; i < 3000; ++i)
        {
            T c[3];
            for (int a = 0; a < 3; ++a)
            {
                double v = lo[a] + len[a] * U(rng);
                switch (int(U(rng) * 12))
                {
                case 0: v = lo[a]; break;
                case 1: v = lo[a] + len[a]; break;
                case 2: v = lo[a] - len[a] * U(rng); break;
                case 3: v = lo[a] + len[a] * (1 + U(rng)); break;
                case 4: v = std::nextafter(lo[a] + len[a], 1e300); break;
                case 5: v = std::nextafter(lo[a], -1e300); break;
                case 6: v = std::numeric_limits<double>::quiet_NaN(); break;
                case 7: v = (i & 1 ? 1 : -1) * std::numeric_limits<double>::infinity(); break;
                default: break;
                }
                c[a] = T(v);
            }
            if (i % 97 == 0) c[0] = c[1] = c[2] = std::numeric_limits<T>::max();
            const Key want = at::binaryKey<T>(c[0], c[1], c[2], g);
            const Key got = at::binaryKeyFast<T>(c[0], c[1], c[2], g, spread);
            ++keys;
            if (want != got && bad++ < 5)
                std::printf("mismatch (%s, box %d): %.17g %.17g %.17g general %016llx fast %016llx\n",
                            sizeof(T) == 4 ? "float" : "double", b, double(c[0]), double(c[1]), double(c[2]),
                            (unsigned long long)want, (unsigned long long)got);
        }
    }
    return bad;
}

int main()
{
    std::mt19937_64 rng(20260101);
    long long boxes = 0, keys = 0;
    const int bad = check<float>(rng, boxes, keys) + check<double>(rng, boxes, keys);
    std::printf("boxes %lld keys %lld mismatches %d\n", boxes, keys, bad);
    return bad ? 1 : 0;
}
