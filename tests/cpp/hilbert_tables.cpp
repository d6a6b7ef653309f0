// Host check of the table-driven Hilbert curve (hilbertKeyTable /
// hilbertDecodeTable, adaptive_trees_amd/csrc/common.hpp) against the bitwise
// levels (hilbertKey / hilbertDecode): equal keys for coordinates whose bits
// cover every octant at every level (all 2^9 points of an 8^3 grid shifted
// through every level, and random 21-bit points), and both decodes invert
// the key. Prints one summary line; exit status 1 on any mismatch. Built and
// run by tests/test_hilbert_tables_cpu.py.
#include <cstdio>
#include <random>

#include "common.hpp"

int main()
{
    using namespace at;
    const auto &T = hilbert_table::kTables;
    long long checked = 0, bad = 0;
    auto check = [&](unsigned x, unsigned y, unsigned z) {
        const Key a = hilbertKey(x, y, z), b = hilbertKeyTable(T.enc, x, y, z);
        unsigned x1, y1, z1, x2, y2, z2;
        hilbertDecode(a, x1, y1, z1);
        hilbertDecodeTable(T.dec, a, x2, y2, z2);
        ++checked;
        if (a != b || x1 != x || y1 != y || z1 != z || x2 != x || y2 != y || z2 != z)
        {
            if (bad++ < 5)
                std::printf("mismatch at (%u, %u, %u): key %llx table %llx\n", x, y, z, (unsigned long long)a,
                            (unsigned long long)b);
        }
    };
    for (int shift = 0; shift + 3 <= 21; ++shift)
        for (unsigned v = 0; v < 512; ++v)
            check((v & 7u) << shift, ((v >> 3) & 7u) << shift, ((v >> 6) & 7u) << shift);
    std::mt19937_64 rng(12345);
    for (int i = 0; i < 2000000; ++i)
    {
        const std::uint64_t r = rng();
        check(unsigned(r) & 0x1fffffu, unsigned(r >> 21) & 0x1fffffu, unsigned(r >> 42) & 0x1fffffu);
    }
    check(0x1fffffu, 0x1fffffu, 0x1fffffu);
    // cell codes by level (both curves): the cell key, a child's key from
    // its parent's code, and the cell's corner from its key
    for (int i = 0; i < 300000; ++i)
    {
        const std::uint64_t r = rng();
        const unsigned x = unsigned(r) & 0x1fffffu, y = unsigned(r >> 21) & 0x1fffffu, z = unsigned(r >> 42) & 0x1fffffu;
        const unsigned level = unsigned(rng() % 22);
        for (const bool hil : {false, true})
        {
            const Key full = sfcEncode(x, y, z, hil);
            const CellCode c = cellCode(x, y, z, level, hil, T.enc);
            bool ok = c.key == enclosingBoxCode(full, level);
            if (level > 0)
            {
                const CellCode p = cellCode(x, y, z, level - 1, hil, T.enc);
                const unsigned b = kOctLevels - level;
                const unsigned raw = ((x >> b) & 1u) << 2 | ((y >> b) & 1u) << 1 | ((z >> b) & 1u);
                ok = ok && childKey(p, raw, level, hil, T.enc) == c.key;
            }
            unsigned cx, cy, cz, st;
            cellCorner(c.key, level, hil, T.dec, cx, cy, cz, st);
            const unsigned keep = level ? ~((1u << (kOctLevels - level)) - 1u) : 0u;
            ok = ok && cx == (x & keep) && cy == (y & keep) && cz == (z & keep) && (!hil || st == c.state);
            ++checked;
            if (!ok && bad++ < 5) std::printf("cell code mismatch at level %u (hilbert %d)\n", level, int(hil));
        }
    }
    std::printf("checked %lld bad %lld\n", checked, bad);
    return bad ? 1 : 0;
}
