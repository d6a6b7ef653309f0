// common.hpp -- shared device/host definitions of the gfx950 adaptive-tree engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace at
{
using Key = std::uint64_t;
constexpr Key kMaxKey = Key(1) << 63;   // root's exclusive end (config.hpp:45-46)
constexpr Key kInvalidKey = ~Key(0);    // escaped-slot key (config.hpp:41)
constexpr unsigned kOctLevels = 21;     // 21 bits per axis, 63-bit keys
constexpr unsigned kBinLevels = 63;     // binary depth cap

// Error with a C-ABI status code attached.
struct Error : std::runtime_error
{
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hipCheck(hipError_t e, const char *what, const char *file, int line);
#define AT_HIP(expr) ::at::hipCheck((expr), #expr, __FILE__, __LINE__)

// Device-resident tree state. The first eight fields are the reference's
// RebalanceState (helpers.cuh:110-120) in the same order, so numLeaves_d(),
// numNodes_d() and numActiveParticles_d() point into it.
struct DevState
{
    int changed;
    int converged;
    int numLeaves;
    int newNumLeaves;
    int needsResize;
    int numNodes;
    int activeBuffer;
    unsigned numActiveParticles;
    // not part of RebalanceState: passes committed since the last reset
    // (an update's passes that changed the tree, before it converged)
    unsigned passes;
};

__host__ __device__ inline int clz64(Key x)
{
#ifdef __HIP_DEVICE_COMPILE__
    return x ? __clzll((long long)x) : 64;
#else
    return x ? __builtin_clzll(x) : 64;
#endif
}
__host__ __device__ inline int clz32(std::uint32_t x)
{
#ifdef __HIP_DEVICE_COMPILE__
    return x ? __clz((int)x) : 32;
#else
    return x ? __builtin_clz(x) : 32;
#endif
}

// ---- octree level helpers: a level-l node spans 2^(3(21-l)) keys ----------
__host__ __device__ inline Key nodeRange(unsigned level) { return Key(1) << (3u * (kOctLevels - level)); }
__host__ __device__ inline unsigned treeLevel(Key range) { return unsigned(clz64(range - 1) - 1) / 3u; }
__host__ __device__ inline unsigned octalDigit(Key code, unsigned pos)
{
    return unsigned(code >> (3u * (kOctLevels - pos))) & 7u;
}
__host__ __device__ inline Key enclosingBoxCode(Key k, unsigned level) { return k & ~(nodeRange(level) - 1); }
// Warren-Salmon placeholder-bit box key of a node with `prefix` path bits.
__host__ __device__ inline Key placeholderKey(Key start, int prefix)
{
    return (Key(1) << prefix) | (start >> (63 - prefix));
}

// ---- Morton: magic-number bit spreading of 21-bit coordinates --------------
__host__ __device__ inline Key spread3(std::uint32_t v)
{
    Key x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}
__host__ __device__ inline std::uint32_t gather3(Key x)
{
    x &= 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
    x = (x ^ (x >> 8)) & 0x1f0000ff0000ffull;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffull;
    x = (x ^ (x >> 32)) & 0x1fffffull;
    return std::uint32_t(x);
}
__host__ __device__ inline Key mortonKey(unsigned x, unsigned y, unsigned z)
{
    return (spread3(x) << 2) | (spread3(y) << 1) | spread3(z);
}

// ---- Hilbert (the reference curve, sfc_keys.cuh:254-331) --------------------
// Per octal digit: visit order of the Morton octant, then reflect and rotate
// the remaining coordinate bits into the child's frame.
__host__ __device__ inline Key hilbertKey(unsigned px, unsigned py, unsigned pz)
{
    // visit order {0,1,3,2,7,6,4,5} packed as 3-bit fields
    constexpr std::uint32_t kVisit = (0u << 0) | (1u << 3) | (3u << 6) | (2u << 9) | (7u << 12) |
                                     (6u << 15) | (4u << 18) | (5u << 21);
    Key key = 0;
#pragma unroll
    for (int level = int(kOctLevels) - 1; level >= 0; --level)
    {
        const unsigned xi = (px >> level) & 1u, yi = (py >> level) & 1u, zi = (pz >> level) & 1u;
        const unsigned nyi = yi ^ 1u, nzi = zi ^ 1u;
        const unsigned oct = (xi << 2) | (yi << 1) | zi;
        key = (key << 3) | Key((kVisit >> (3 * oct)) & 7u);
        px ^= 0u - (xi & (nyi | zi));
        py ^= 0u - ((xi & (yi | zi)) | (yi & nzi));
        pz ^= 0u - ((xi & nyi & nzi) | (yi & nzi));
        const unsigned ox = px, oy = py, oz = pz;
        if (zi) { px = oy; py = oz; pz = ox; }
        else if (!yi) { px = oz; pz = ox; }
    }
    return key;
}
__host__ __device__ inline void hilbertDecode(Key key, unsigned &ox, unsigned &oy, unsigned &oz)
{
    unsigned px = 0, py = 0, pz = 0;
#pragma unroll
    for (unsigned level = 0; level < kOctLevels; ++level)
    {
        const unsigned oct = unsigned(key >> (3 * level)) & 7u;
        const unsigned xi = oct >> 2, yi = (oct >> 1) & 1u, zi = oct & 1u;
        const unsigned nxi = xi ^ 1u, nyi = yi ^ 1u, nzi = zi ^ 1u;
        const unsigned tx = px, ty = py, tz = pz;
        if (yi ^ zi) { px = tz; pz = ty; py = tx; }
        else if ((nxi & nyi & nzi) | (xi & yi & zi)) { px = tz; pz = tx; }
        const unsigned mask = (1u << level) - 1u;
        px ^= mask & (0u - (xi & (yi | zi)));
        py ^= mask & (0u - ((xi & (nyi | nzi)) | (nxi & yi & zi)));
        pz ^= mask & (0u - ((xi & nyi & nzi) | (yi & zi)));
        px |= xi << level;
        py |= (xi ^ yi) << level;
        pz |= (yi ^ zi) << level;
    }
    ox = px; oy = py; oz = pz;
}
// ---- Hilbert by table: the same curve as a 48-state automaton ---------------
// Every level of hilbertKey maps the remaining coordinate words through a
// signed permutation (which word becomes x, y, z, and which are
// complemented), so the map accumulated over the levels above is one of 48
// states: state = 8 * (index of the axis permutation) + flip mask. One table
// row per (state, octant) gives the level's digit and the next state:
//   kHilbertEnc[8 s + raw octant of the input bits]     = digit | next << 3
//   kHilbertDec[8 s + digit]                            = raw octant | next << 3
// (raw octant = x << 2 | y << 1 | z of the untransformed coordinates). The
// tables are derived at compile time from the bitwise levels above;
// tests/cpp/hilbert_tables.cpp checks them against hilbertKey/hilbertDecode.
namespace hilbert_table
{
constexpr int kPerms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
constexpr int permIndex(int a, int b, int c)
{
    for (int i = 0; i < 6; ++i)
        if (kPerms[i][0] == a && kPerms[i][1] == b && kPerms[i][2] == c) return i;
    return -1;
}
// The transition of state s for transformed bits (xi, yi, zi) = oct.
constexpr int next(int s, unsigned oct)
{
    const int *p = kPerms[s >> 3];
    int f[3] = {(s >> 2) & 1, (s >> 1) & 1, s & 1}; // flip of output word 0, 1, 2
    const unsigned xi = oct >> 2, yi = (oct >> 1) & 1u, zi = oct & 1u;
    const unsigned nyi = yi ^ 1u, nzi = zi ^ 1u;
    f[0] ^= int(xi & (nyi | zi));
    f[1] ^= int((xi & (yi | zi)) | (yi & nzi));
    f[2] ^= int((xi & nyi & nzi) | (yi & nzi));
    int q[3] = {p[0], p[1], p[2]}, g[3] = {f[0], f[1], f[2]};
    if (zi)
    {
        q[0] = p[1], q[1] = p[2], q[2] = p[0];
        g[0] = f[1], g[1] = f[2], g[2] = f[0];
    }
    else if (!yi)
    {
        q[0] = p[2], q[2] = p[0];
        g[0] = f[2], g[2] = f[0];
    }
    return permIndex(q[0], q[1], q[2]) * 8 + (g[0] << 2 | g[1] << 1 | g[2]);
}
// Transformed octant of state s for the raw input bits r (x << 2 | y << 1 | z).
constexpr unsigned transformed(int s, unsigned r)
{
    const int *p = kPerms[s >> 3];
    unsigned o = 0;
    for (int i = 0; i < 3; ++i)
    {
        const unsigned bit = ((r >> (2 - p[i])) & 1u) ^ unsigned((s >> (2 - i)) & 1);
        o |= bit << (2 - i);
    }
    return o;
}
constexpr unsigned kVisit[8] = {0, 1, 3, 2, 7, 6, 4, 5};
struct Tables
{
    std::uint16_t enc[48 * 8];
    std::uint16_t dec[48 * 8];
};
constexpr Tables make()
{
    Tables t{};
    for (int s = 0; s < 48; ++s)
        for (unsigned r = 0; r < 8; ++r)
        {
            const unsigned oct = transformed(s, r);
            const unsigned digit = kVisit[oct];
            const unsigned nx = unsigned(next(s, oct));
            t.enc[8 * s + int(r)] = std::uint16_t(digit | nx << 3);
            t.dec[8 * s + int(digit)] = std::uint16_t(r | nx << 3);
        }
    return t;
}
inline constexpr Tables kTables = make();
constexpr int kStart = 0; // identity permutation, nothing complemented
} // namespace hilbert_table

// Encode / decode through the tables (enc / dec: kTables.enc / .dec, e.g.
// staged in LDS by the calling kernel).
__host__ __device__ inline Key hilbertKeyTable(const std::uint16_t *enc, unsigned px, unsigned py, unsigned pz)
{
    Key key = 0;
    unsigned s = hilbert_table::kStart;
#pragma unroll
    for (int level = int(kOctLevels) - 1; level >= 0; --level)
    {
        const unsigned r = ((px >> level) & 1u) << 2 | ((py >> level) & 1u) << 1 | ((pz >> level) & 1u);
        const unsigned e = enc[8 * s + r];
        key = (key << 3) | Key(e & 7u);
        s = e >> 3;
    }
    return key;
}
__host__ __device__ inline void hilbertDecodeTable(const std::uint16_t *dec, Key key, unsigned &ox, unsigned &oy,
                                                   unsigned &oz)
{
    unsigned x = 0, y = 0, z = 0, s = hilbert_table::kStart;
#pragma unroll
    for (int level = int(kOctLevels) - 1; level >= 0; --level)
    {
        const unsigned d = unsigned(key >> (3 * level)) & 7u;
        const unsigned e = dec[8 * s + d];
        x = (x << 1) | ((e >> 2) & 1u);
        y = (y << 1) | ((e >> 1) & 1u);
        z = (z << 1) | (e & 1u);
        s = e >> 3;
    }
    ox = x, oy = y, oz = z;
}

// Cells by their level: the key of the level-`level` cell holding grid
// point (x, y, z) (its lower digits zero) and, for Hilbert, the table state
// below it; a child cell's key from its parent's code and the child's raw
// octant (the coordinate bits at the child's level); the corner of a cell
// from its key. Hilbert codes only walk the `level` digits above.
struct CellCode
{
    Key key;
    unsigned state;
};
__host__ __device__ inline CellCode cellCode(unsigned x, unsigned y, unsigned z, unsigned level, bool hilbert,
                                             const std::uint16_t *enc)
{
    if (!hilbert) return CellCode{mortonKey(x, y, z) & ~(nodeRange(level) - 1), 0u};
    Key key = 0;
    unsigned s = hilbert_table::kStart;
    for (unsigned l = 0; l < level; ++l)
    {
        const unsigned b = kOctLevels - 1 - l;
        const unsigned r = ((x >> b) & 1u) << 2 | ((y >> b) & 1u) << 1 | ((z >> b) & 1u);
        const unsigned e = enc[8 * s + r];
        key = (key << 3) | Key(e & 7u);
        s = e >> 3;
    }
    return CellCode{level ? key << (3u * (kOctLevels - level)) : Key(0), s};
}
__host__ __device__ inline Key childKey(const CellCode &parent, unsigned raw, unsigned level, bool hilbert,
                                        const std::uint16_t *enc)
{
    const unsigned digit = hilbert ? unsigned(enc[8 * parent.state + raw]) & 7u : raw;
    return parent.key | (Key(digit) << (3u * (kOctLevels - level)));
}
// Corner of the level-`level` cell with key `key` (coordinate bits below
// the level zero) and its Hilbert state.
__host__ __device__ inline void cellCorner(Key key, unsigned level, bool hilbert, const std::uint16_t *dec,
                                           unsigned &x, unsigned &y, unsigned &z, unsigned &state)
{
    const unsigned low = kOctLevels - level;
    if (!hilbert)
    {
        const std::uint32_t keep = ~((1u << low) - 1u);
        x = gather3(key >> 2) & keep, y = gather3(key >> 1) & keep, z = gather3(key) & keep;
        state = 0;
        return;
    }
    unsigned cx = 0, cy = 0, cz = 0, s = hilbert_table::kStart;
    for (unsigned l = 0; l < level; ++l)
    {
        const unsigned d = unsigned(key >> (3u * (kOctLevels - 1 - l))) & 7u;
        const unsigned e = dec[8 * s + d];
        cx = (cx << 1) | ((e >> 2) & 1u);
        cy = (cy << 1) | ((e >> 1) & 1u);
        cz = (cz << 1) | (e & 1u);
        s = e >> 3;
    }
    x = level ? cx << low : 0u, y = level ? cy << low : 0u, z = level ? cz << low : 0u;
    state = s;
}

__host__ __device__ inline Key sfcEncode(unsigned x, unsigned y, unsigned z, bool hilbert)
{
    return hilbert ? hilbertKey(x, y, z) : mortonKey(x, y, z);
}
__host__ __device__ inline void sfcDecode(Key k, bool hilbert, unsigned &x, unsigned &y, unsigned &z)
{
    if (hilbert) hilbertDecode(k, x, y, z);
    else { x = gather3(k >> 2); y = gather3(k >> 1); z = gather3(k); }
}
// The same through the Hilbert tables (enc / dec as in hilbertKeyTable).
__host__ __device__ inline Key sfcEncodeT(unsigned x, unsigned y, unsigned z, bool hilbert, const std::uint16_t *enc)
{
    return hilbert ? hilbertKeyTable(enc, x, y, z) : mortonKey(x, y, z);
}
__host__ __device__ inline void sfcDecodeT(Key k, bool hilbert, const std::uint16_t *dec, unsigned &x, unsigned &y,
                                           unsigned &z)
{
    if (hilbert) hilbertDecodeTable(dec, k, x, y, z);
    else { x = gather3(k >> 2); y = gather3(k >> 1); z = gather3(k); }
}

// ---- float box -> integer grid (sfc_keys.cuh:420-496) -------------------------
template <class T>
struct GridBox
{
    T xmin, ymin, zmin;
    T mx, my, mz;          // 2^bits * (1/length), computed on the host in T
    int cx, cy, cz;        // 2^bits per axis
};

template <class T>
__host__ __device__ inline bool isEscaped(T x, T y, T z);
template <>
__host__ __device__ inline bool isEscaped<float>(float x, float y, float z)
{
    constexpr float e = 3.40282346638528859811704183484516925e+38f;
    return x == e && y == e && z == e;
}
template <>
__host__ __device__ inline bool isEscaped<double>(double x, double y, double z)
{
    constexpr double e = 1.79769313486231570814527423731704357e+308;
    return x == e && y == e && z == e;
}

// float/double -> int with saturation (NaN -> 0): the device conversion
// instruction's own semantics, spelled out so the host path agrees.
template <class T>
__host__ __device__ inline int toIntSat(T v)
{
    if (!(v == v)) return 0;
    if (v >= T(2147483648.0)) return 2147483647;
    if (v <= T(-2147483648.0)) return -2147483647 - 1;
    return int(v);
}

template <class T>
__host__ __device__ inline void gridCell(T x, T y, T z, const GridBox<T> &b, unsigned &cx, unsigned &cy, unsigned &cz)
{
    int ix = toIntSat<T>(fma(-b.xmin, b.mx, floor(x * b.mx)));
    int iy = toIntSat<T>(fma(-b.ymin, b.my, floor(y * b.my)));
    int iz = toIntSat<T>(fma(-b.zmin, b.mz, floor(z * b.mz)));
    ix = ix < b.cx - 1 ? ix : b.cx - 1;
    iy = iy < b.cy - 1 ? iy : b.cy - 1;
    iz = iz < b.cz - 1 ? iz : b.cz - 1;
    cx = unsigned(ix), cy = unsigned(iy), cz = unsigned(iz);
}
template <class T>
__host__ __device__ inline Key gridKey(T x, T y, T z, const GridBox<T> &b, bool hilbert)
{
    unsigned cx, cy, cz;
    gridCell<T>(x, y, z, b, cx, cy, cz);
    return sfcEncode(cx, cy, cz, hilbert);
}
// The same key through the Hilbert tables (Hilbert trees only).
template <class T>
__host__ __device__ inline Key gridKeyHilbertTable(T x, T y, T z, const GridBox<T> &b, const std::uint16_t *enc)
{
    unsigned cx, cy, cz;
    gridCell<T>(x, y, z, b, cx, cy, cz);
    return hilbertKeyTable(enc, cx, cy, cz);
}

// ---- binary longest-side keys (binary_node.cuh:43-59) ------------------------

// min(floor(clamp(v, 0, 1) * 2^63), 2^63 - 1). The product is exact, so the
// integer part is taken from the bits: c = m 2^(e - 1075) with the 53-bit
// significand m gives m << (e - 1012) or m >> (1012 - e). Zero, subnormals
// and NaN give 0 (NaN as the device's double -> integer conversion does).
__host__ __device__ inline Key fixed63(double v)
{
    const double c1 = (1.0 < v) ? 1.0 : v;
    const double c0 = (0.0 > c1) ? 0.0 : c1;
    std::uint64_t bits;
    __builtin_memcpy(&bits, &c0, sizeof(bits));
    const int e = int((bits >> 52) & 0x7ff);
    const Key m = (bits & ((Key(1) << 52) - 1)) | (Key(1) << 52);
    const int sh = e - 1012;
    Key q;
    if (e == 0 || e == 0x7ff) q = 0;
    else if (sh >= 0) q = m << sh; // sh <= 11 (c0 <= 1)
    else q = sh <= -64 ? Key(0) : m >> -sh;
    return q < kMaxKey - 1 ? q : kMaxKey - 1;
}

struct BinaryGeom
{
    double lo[3];       // box minimum, widened from Real
    double len[3];      // box length, computed in double from the Real bounds
    std::uint8_t axis[kBinLevels];
    // Fast path: the schedule is a short prefix followed by a period-3 cycle
    // of three distinct axes; the cycle is a 3-way bit interleave.
    int periodic;       // 1 if the decomposition below describes axis[]
    int prefix;         // number of leading depths handled one by one
    int consumed[3];    // bits each axis gave to the prefix
    int order[3];       // axes of one cycle, most significant first
    int triples;        // ceil((63 - prefix) / 3)
    int prefixAxis;     // the only axis of the prefix levels, or -1 if they mix axes
    double inv[3];      // correctly rounded 1 / len
    int fastDiv;        // 1: every len and inv lies in [2^-1000, 2^1000] (binaryUnit's fast form)
    // Branch-free form (binaryKeyFast): fastDiv, a periodic schedule whose
    // prefix levels all split one axis, and at most 32 key bits per axis.
    // Slot i of the cycle splits axis order[i] and contributes the top n_i
    // bits of that axis's 63-bit fixed point: n_i = triples, plus prefix for
    // the prefix axis.
    int fast;
    double slotLo[3], slotLen[3], slotInv[3]; // lo, len, inv of axis order[i]
    double slotScale[3];      // 2^n_i
    std::uint32_t slotMax[3]; // 2^n_i - 1
};

// (x - lo) / len, as far as fixed63 can tell: the correctly rounded quotient
// on (0, len), and 0 / 1 below / above it. The fast form is Markstein's
// division (IBM J. Res. Develop. 34(1), 1990): from q = a * inv a residual
// r = fma(-len, q, a) is exact, and q + r * inv rounded once is the
// correctly rounded quotient when inv = RN(1 / len) and q is within one ulp
// of a / len; the first correction brings q within one ulp, the second
// rounds it. The operands stay far from overflow and underflow by fastDiv
// and 0 < a < len.
__host__ __device__ inline double binaryUnit(double a, double len, double inv, int fast)
{
    if (!fast) return a / len;
    if (a <= 0.0) return 0.0;
    if (a >= len) return 1.0;
    double q = a * inv;
    double r = fma(-len, q, a);
    q = fma(r, inv, q);
    r = fma(-len, q, a);
    return fma(r, inv, q);
}

// Longest-side key of normalised coordinates v (binary_node.cuh:43-59):
// depth d takes the next most significant bit of axis[d]'s 63-bit fixed point.
__host__ __device__ inline Key binaryInterleave(Key f0, Key f1, Key f2, const BinaryGeom &g)
{
    Key f[3] = {f0, f1, f2};
    if (g.periodic)
    {
        Key hi = 0;
        if (g.prefixAxis >= 0)
        {
            // the prefix is the top `prefix` bits of one axis
            if (g.prefix > 0)
            {
                const Key lowBits = (Key(1) << g.prefix) - 1;
                if (g.prefixAxis == 0) hi = (f0 >> (63 - g.prefix)) & lowBits, f0 <<= g.prefix;
                else if (g.prefixAxis == 1) hi = (f1 >> (63 - g.prefix)) & lowBits, f1 <<= g.prefix;
                else hi = (f2 >> (63 - g.prefix)) & lowBits, f2 <<= g.prefix;
            }
        }
        else
        for (int d = 0; d < g.prefix; ++d)
        {
            const int a = g.axis[d];
            const Key fa = a == 0 ? f0 : (a == 1 ? f1 : f2);
            hi = (hi << 1) | ((fa >> 62) & 1u);
            if (a == 0) f0 <<= 1; else if (a == 1) f1 <<= 1; else f2 <<= 1;
        }
        const int R = 63 - g.prefix;
        if (R == 0) return hi;
        const int m = g.triples;
        const Key fr[3] = {f0, f1, f2};
        const Key mask = (Key(1) << m) - 1;
        const Key t0 = (fr[g.order[0]] >> (63 - m)) & mask;
        const Key t1 = (fr[g.order[1]] >> (63 - m)) & mask;
        const Key t2 = (fr[g.order[2]] >> (63 - m)) & mask;
        const Key full = (spread3(std::uint32_t(t0)) << 2) | (spread3(std::uint32_t(t1)) << 1) |
                         spread3(std::uint32_t(t2));
        const Key low = full >> (3 * m - R);
        return (hi << R) | low;
    }
    Key key = 0;
    for (unsigned d = 0; d < kBinLevels; ++d)
    {
        Key &c = f[g.axis[d]];
        key = (key << 1) | ((c >> 62) & 1u);
        c <<= 1;
    }
    return key;
}

// binaryKey for geometries with g.fast, without branches. Per slot: the
// quotient q as binaryUnit computes it (selects for its two early returns),
// then the top n bits of fixed63(q) directly as floor(q * 2^n) -- exact, as
// the scaling is by a power of two -- with 2^n - 1 at q = 1 (fixed63 clamps
// to 2^63 - 2 there) and 0 for NaN. The prefix bits are the bits above the
// low `triples` of the prefix slot (the other slots have none), and the
// three slots are interleaved a byte at a time through `spread`, the
// 256-entry table of spread3.
template <class T>
__host__ __device__ inline Key binaryKeyFast(T x, T y, T z, const BinaryGeom &g, const std::uint32_t *spread)
{
    std::uint32_t t[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        const int a = g.order[i];
        const double av = double(a == 0 ? x : (a == 1 ? y : z)) - g.slotLo[i];
        const double len = g.slotLen[i], inv = g.slotInv[i];
        double q = av * inv;
        double r = fma(-len, q, av);
        q = fma(r, inv, q);
        r = fma(-len, q, av);
        q = fma(r, inv, q);
        q = av <= 0.0 ? 0.0 : q;
        q = av >= len ? 1.0 : q;
        const double s = q * g.slotScale[i];
        const bool one = s >= g.slotScale[i];
        const double sc = s >= 0.0 && !one ? s : 0.0; // NaN -> 0
        t[i] = one ? g.slotMax[i] : std::uint32_t(sc);
    }
    const int m = g.triples;
    const std::uint32_t hi = (t[0] | t[1] | t[2]) >> m;
    const std::uint32_t mask = (std::uint32_t(1) << m) - 1;
#pragma unroll
    for (int i = 0; i < 3; ++i) t[i] &= mask;
    auto bytes = [&](int sh) {
        return (spread[(t[0] >> sh) & 255u] << 2) | (spread[(t[1] >> sh) & 255u] << 1) | spread[(t[2] >> sh) & 255u];
    };
    const Key full = Key(bytes(0)) | (Key(bytes(8)) << 24) | (Key(bytes(16)) << 48);
    const int R = int(kBinLevels) - g.prefix;
    const Key key = (Key(hi) << R) | (full >> (3 * m - R));
    return isEscaped<T>(x, y, z) ? kInvalidKey : key;
}

template <class T>
__host__ __device__ inline Key binaryKey(T x, T y, T z, const BinaryGeom &g)
{
    if (isEscaped<T>(x, y, z)) return kInvalidKey;
    const double v0 = binaryUnit(double(x) - g.lo[0], g.len[0], g.inv[0], g.fastDiv);
    const double v1 = binaryUnit(double(y) - g.lo[1], g.len[1], g.inv[1], g.fastDiv);
    const double v2 = binaryUnit(double(z) - g.lo[2], g.len[2], g.inv[2], g.fastDiv);
    return binaryInterleave(fixed63(v0), fixed63(v1), fixed63(v2), g);
}

// ---- integer sequences -------------------------------------------------------
template <class T>
__host__ __device__ inline std::size_t lowerBound(const T *a, std::size_t n, T v)
{
    std::size_t first = 0;
    while (n > 0)
    {
        const std::size_t half = n >> 1;
        if (a[first + half] < v) { first += half + 1; n -= half + 1; }
        else n = half;
    }
    return first;
}
template <class T>
__host__ __device__ inline std::size_t upperBound(const T *a, std::size_t n, T v)
{
    std::size_t first = 0;
    while (n > 0)
    {
        const std::size_t half = n >> 1;
        if (!(v < a[first + half])) { first += half + 1; n -= half + 1; }
        else n = half;
    }
    return first;
}

inline unsigned ceilDiv(std::size_t n, unsigned b) { return unsigned((n + b - 1) / b); }

// ---- sampled search index of the sorted keys -----------------------------------
// Q[j] = P[j * kIndexStride]. A lower bound first searches Q (small enough to
// stay in L2 / MALL), then one stride window of P, so a search over tens of
// millions of keys touches only a few lines of P. Two searches run in lock
// step (their loads overlap); the branch-free form has a data-independent
// step count. Results equal lowerBound over P.
constexpr std::size_t kIndexStride = 128;
__host__ __device__ inline std::size_t indexSize(std::size_t n) { return (n + kIndexStride - 1) / kIndexStride; }

__device__ inline void lowerBoundPair(const Key *a, std::size_t n, Key v0, Key v1, std::size_t &r0, std::size_t &r1)
{
    if (n == 0)
    {
        r0 = r1 = 0;
        return;
    }
    std::size_t b0 = 0, b1 = 0;
    while (n > 1)
    {
        const std::size_t half = n >> 1;
        const Key x0 = a[b0 + half], x1 = a[b1 + half];
        b0 = x0 < v0 ? b0 + half : b0;
        b1 = x1 < v1 ? b1 + half : b1;
        n -= half;
    }
    r0 = b0 + (a[b0] < v0 ? 1 : 0);
    r1 = b1 + (a[b1] < v1 ? 1 : 0);
}

// The answers r = lower_bound(a, n, v) given k = lower_bound over the
// samples a[0], a[S], a[2S], ... (ceil(n/S) of them): r lies in
// ((k-1) S, k S] (clipped to n), found by branch-free halving inside that
// window; two searches in lock step.
__device__ inline void refinePair(const Key *a, std::size_t n, std::size_t S, std::size_t k0, std::size_t k1, Key v0,
                                  Key v1, std::size_t &r0, std::size_t &r1)
{
    const std::size_t m = (n + S - 1) / S;
    std::size_t b0 = k0 ? (k0 - 1) * S + 1 : 0, b1 = k1 ? (k1 - 1) * S + 1 : 0;
    std::size_t l0 = k0 ? (k0 < m ? k0 * S : n) - b0 : 0, l1 = k1 ? (k1 < m ? k1 * S : n) - b1 : 0;
    while (l0 > 1 || l1 > 1)
    {
        if (l0 > 1)
        {
            const std::size_t h = l0 >> 1;
            b0 = a[b0 + h] < v0 ? b0 + h : b0;
            l0 -= h;
        }
        if (l1 > 1)
        {
            const std::size_t h = l1 >> 1;
            b1 = a[b1 + h] < v1 ? b1 + h : b1;
            l1 -= h;
        }
    }
    r0 = l0 ? b0 + (a[b0] < v0 ? 1 : 0) : b0;
    r1 = l1 ? b1 + (a[b1] < v1 ? 1 : 0) : b1;
}

// lower_bound(P, n, v0) and lower_bound(P, n, v1) through the index Q.
__device__ inline void lowerBoundIndexedPair(const Key *P, std::size_t n, const Key *Q, Key v0, Key v1,
                                             std::size_t &r0, std::size_t &r1)
{
    std::size_t j0, j1;
    lowerBoundPair(Q, indexSize(n), v0, v1, j0, j1);
    refinePair(P, n, kIndexStride, j0, j1, v0, v1, r0, r1);
}

// Lower bounds of v0 and v1 inside P[lo, hi) when both lie in [lo, hi]:
// lo + the number of keys of P[lo, hi) below each value.
__device__ inline void lowerBoundPairIn(const Key *P, std::size_t lo, std::size_t hi, Key v0, Key v1,
                                        std::size_t &r0, std::size_t &r1)
{
    std::size_t n = hi - lo;
    if (n == 0)
    {
        r0 = r1 = lo;
        return;
    }
    std::size_t b0 = lo, b1 = lo;
    while (n > 1)
    {
        const std::size_t half = n >> 1;
        const Key x0 = P[b0 + half], x1 = P[b1 + half];
        b0 = x0 < v0 ? b0 + half : b0;
        b1 = x1 < v1 ? b1 + half : b1;
        n -= half;
    }
    r0 = b0 + (P[b0] < v0 ? 1 : 0);
    r1 = b1 + (P[b1] < v1 ? 1 : 0);
}

// 32-bit index forms of refinePair and lowerBoundPairIn (arrays of fewer than
// 2^32 keys): the same steps in half the registers.
__device__ inline void refinePair32(const Key *a, std::uint32_t n, std::uint32_t S, std::uint32_t k0,
                                    std::uint32_t k1, Key v0, Key v1, std::uint32_t &r0, std::uint32_t &r1)
{
    const std::uint32_t m = std::uint32_t((std::uint64_t(n) + S - 1) / S);
    std::uint32_t b0 = k0 ? (k0 - 1) * S + 1 : 0, b1 = k1 ? (k1 - 1) * S + 1 : 0;
    std::uint32_t l0 = k0 ? (k0 < m ? k0 * S : n) - b0 : 0, l1 = k1 ? (k1 < m ? k1 * S : n) - b1 : 0;
    while (l0 > 1 || l1 > 1)
    {
        if (l0 > 1)
        {
            const std::uint32_t h = l0 >> 1;
            b0 = a[b0 + h] < v0 ? b0 + h : b0;
            l0 -= h;
        }
        if (l1 > 1)
        {
            const std::uint32_t h = l1 >> 1;
            b1 = a[b1 + h] < v1 ? b1 + h : b1;
            l1 -= h;
        }
    }
    r0 = l0 ? b0 + (a[b0] < v0 ? 1u : 0u) : b0;
    r1 = l1 ? b1 + (a[b1] < v1 ? 1u : 0u) : b1;
}

__device__ inline void lowerBoundPairIn32(const Key *P, std::uint32_t lo, std::uint32_t hi, Key v0, Key v1,
                                          std::uint32_t &r0, std::uint32_t &r1)
{
    std::uint32_t n = hi - lo;
    if (n == 0)
    {
        r0 = r1 = lo;
        return;
    }
    std::uint32_t b0 = lo, b1 = lo;
    while (n > 1)
    {
        const std::uint32_t half = n >> 1;
        const Key x0 = P[b0 + half], x1 = P[b1 + half];
        b0 = x0 < v0 ? b0 + half : b0;
        b1 = x1 < v1 ? b1 + half : b1;
        n -= half;
    }
    r0 = b0 + (P[b0] < v0 ? 1u : 0u);
    r1 = b1 + (P[b1] < v1 ? 1u : 0u);
}

// ---- device-scope atomics used by the chained scans ---------------------------
__device__ inline std::uint32_t ldAcq(const std::uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void stRel(std::uint32_t *p, std::uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline std::uint64_t ldAcq64(const std::uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void stRel64(std::uint64_t *p, std::uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
} // namespace at
