// oracle.cpp -- CPU restatement of the reference adaptive-tree build/update
// path (nrottstegge/Adaptive_Trees v3).
//
// TEST INFRASTRUCTURE ONLY: the parity checker of tests/ and the CPU
// baseline of bench.py. Nothing in the product library links or calls it.
//
// Each function states the reference lines it restates. The restatement keeps
// the reference's semantics step by step (Jacobi passes, dirty-flag caching,
// capacity rejection, A/B buffers, thrust capacity growth) but is written as
// plain sequential/OpenMP C++ over host arrays.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc
{
using Key = std::uint64_t;
constexpr Key kMaxKey = Key(1) << 63;           // config.hpp:45-46
constexpr Key kInvalidKey = ~Key(0);            // config.hpp:41
constexpr unsigned kOctLevels = 21;             // maxTreeLevel<u64> (sfc_keys.cuh:29-33)
constexpr unsigned kBinLevels = 63;             // maxBinaryDepth (config.hpp:44)

thread_local std::string g_error;

struct Error : std::runtime_error
{
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline int clz64(Key x) { return x ? __builtin_clzll(x) : 64; }
inline int clz32(std::uint32_t x) { return x ? __builtin_clz(x) : 32; }

// ---------------------------------------------------------------------------
// Octree level helpers (sfc_keys.cuh:92-142)
// ---------------------------------------------------------------------------
inline Key nodeRange(unsigned level) { return Key(1) << (3u * (kOctLevels - level)); }
inline unsigned treeLevel(Key range) { return unsigned(clz64(range - 1) - 1) / 3u; }
inline unsigned octalDigit(Key code, unsigned pos)
{
    return unsigned(code >> (3u * (kOctLevels - pos))) & 7u;
}
// log8ceil instantiated on a 32-bit operand (maxTreeLevel<unsigned> = 10,
// unusedBits<unsigned> = 2), as processNode calls it (octree_rebalance.cuh:344).
inline unsigned log8ceil32(std::uint32_t n)
{
    if (n == 0) return 0;
    return 10u - unsigned(clz32(n - 1) - 2) / 3u;
}
inline Key enclosingBoxCode(Key k, unsigned level) { return k & ~(nodeRange(level) - 1); }
inline Key placeholder(Key code, int prefix)
{
    return (Key(1) << prefix) | (code >> (63 - prefix));
}

// ---------------------------------------------------------------------------
// Morton / Hilbert (sfc_keys.cuh:174-331). Morton is written as an explicit
// bit loop (x most significant within each octal digit).
// ---------------------------------------------------------------------------
Key morton(unsigned x, unsigned y, unsigned z)
{
    Key k = 0;
    for (int b = int(kOctLevels) - 1; b >= 0; --b)
        k = (k << 3) | (Key((x >> b) & 1u) << 2) | (Key((y >> b) & 1u) << 1) | Key((z >> b) & 1u);
    return k;
}
void mortonDecode(Key k, unsigned *xyz)
{
    unsigned x = 0, y = 0, z = 0;
    for (unsigned b = 0; b < kOctLevels; ++b)
    {
        z |= unsigned((k >> (3 * b)) & 1u) << b;
        y |= unsigned((k >> (3 * b + 1)) & 1u) << b;
        x |= unsigned((k >> (3 * b + 2)) & 1u) << b;
    }
    xyz[0] = x; xyz[1] = y; xyz[2] = z;
}
// iHilbert (sfc_keys.cuh:254-290): octant visiting order and the per-level
// reflect/rotate of the remaining coordinate bits.
Key hilbert(unsigned px, unsigned py, unsigned pz)
{
    static const unsigned char visit[8] = {0, 1, 3, 2, 7, 6, 4, 5};
    Key key = 0;
    for (int level = int(kOctLevels) - 1; level >= 0; --level)
    {
        const unsigned xi = (px >> level) & 1u, yi = (py >> level) & 1u, zi = (pz >> level) & 1u;
        const unsigned nyi = yi ^ 1u, nzi = zi ^ 1u;
        key = (key << 3) + visit[(xi << 2) | (yi << 1) | zi];
        px ^= 0u - (xi & (nyi | zi));
        py ^= 0u - ((xi & (yi | zi)) | (yi & nzi));
        pz ^= 0u - ((xi & nyi & nzi) | (yi & nzi));
        if (zi) { const unsigned t = px; px = py; py = pz; pz = t; }
        else if (!yi) { const unsigned t = px; px = pz; pz = t; }
    }
    return key;
}
// decodeHilbert (sfc_keys.cuh:293-331)
void hilbertDecode(Key key, unsigned *xyz)
{
    unsigned px = 0, py = 0, pz = 0;
    for (unsigned level = 0; level < kOctLevels; ++level)
    {
        const unsigned oct = unsigned(key >> (3 * level)) & 7u;
        const unsigned xi = oct >> 2, yi = (oct >> 1) & 1u, zi = oct & 1u;
        const unsigned nxi = xi ^ 1u, nyi = yi ^ 1u, nzi = zi ^ 1u;
        if (yi ^ zi) { const unsigned t = px; px = pz; pz = py; py = t; }
        else if ((nxi & nyi & nzi) || (xi & yi & zi)) { const unsigned t = px; px = pz; pz = t; }
        const unsigned mask = (1u << level) - 1u;
        px ^= mask & (0u - (xi & (yi | zi)));
        py ^= mask & (0u - ((xi & (nyi | nzi)) | (nxi & yi & zi)));
        pz ^= mask & (0u - ((xi & nyi & nzi) | (yi & zi)));
        px |= xi << level;
        py |= (xi ^ yi) << level;
        pz |= (yi ^ zi) << level;
    }
    xyz[0] = px; xyz[1] = py; xyz[2] = pz;
}
inline Key sfcEncode(unsigned x, unsigned y, unsigned z, bool h) { return h ? hilbert(x, y, z) : morton(x, y, z); }
inline void sfcDecode(Key k, bool h, unsigned *xyz) { h ? hilbertDecode(k, xyz) : mortonDecode(k, xyz); }

// ---------------------------------------------------------------------------
// Floating-point box -> integer grid (sfc_keys.cuh:420-496)
// ---------------------------------------------------------------------------
template <class T>
struct FBox
{
    T xmin, ymin, zmin, ilx, ily, ilz;
    unsigned bx, by, bz; // bits per axis
};
template <class T>
FBox<T> makeBox(const at_box &b)
{
    FBox<T> r;
    const T xmin = T(b.xmin), xmax = T(b.xmax), ymin = T(b.ymin), ymax = T(b.ymax),
            zmin = T(b.zmin), zmax = T(b.zmax);
    r.xmin = xmin; r.ymin = ymin; r.zmin = zmin;
    r.ilx = T(1) / (xmax - xmin);
    r.ily = T(1) / (ymax - ymin);
    r.ilz = T(1) / (zmax - zmin);
    const T lx = xmax - xmin, ly = ymax - ymin, lz = zmax - zmin;
    const T mxy = lx > ly ? lx : ly;
    const T maxDim = mxy > lz ? mxy : lz;
    const T bias = T(0.5849625007211563);
    r.bx = kOctLevels - static_cast<unsigned>(std::floor(std::log2(maxDim / lx) + bias));
    r.by = kOctLevels - static_cast<unsigned>(std::floor(std::log2(maxDim / ly) + bias));
    r.bz = kOctLevels - static_cast<unsigned>(std::floor(std::log2(maxDim / lz) + bias));
    return r;
}
// GPU float->int conversion saturates and maps NaN to 0; this models it so
// the oracle is defined for every input the device path accepts.
template <class T>
inline int toIntSat(T v)
{
    if (v != v) return 0;
    if (v >= T(2147483648.0)) return std::numeric_limits<int>::max();
    if (v <= T(-2147483648.0)) return std::numeric_limits<int>::min();
    return int(v);
}
// sfc3D (sfc_keys.cuh:468-496). The subtraction of xmin*m is written as one
// fused multiply-add: the device compiler contracts floor(x*m) - xmin*m the
// same way. With xmin == 0 (every configured workload) it is exact either way.
template <class T>
Key sfc3D(T x, T y, T z, const FBox<T> &b, bool h)
{
    const unsigned cx = 1u << b.bx, cy = 1u << b.by, cz = 1u << b.bz;
    const T mx = T(cx) * b.ilx, my = T(cy) * b.ily, mz = T(cz) * b.ilz;
    int ix = toIntSat<T>(std::fma(-b.xmin, mx, std::floor(x * mx)));
    int iy = toIntSat<T>(std::fma(-b.ymin, my, std::floor(y * my)));
    int iz = toIntSat<T>(std::fma(-b.zmin, mz, std::floor(z * mz)));
    ix = std::min(ix, int(cx) - 1);
    iy = std::min(iy, int(cy) - 1);
    iz = std::min(iz, int(cz) - 1);
    return sfcEncode(unsigned(ix), unsigned(iy), unsigned(iz), h);
}
template <class T>
inline bool isEscaped(T x, T y, T z) // config.hpp:103-113
{
    const T e = std::numeric_limits<T>::max();
    return x == e && y == e && z == e;
}

// ---------------------------------------------------------------------------
// Binary (longest-side) keys (binary_node.cuh:12-59)
// ---------------------------------------------------------------------------
struct Geometry { std::uint8_t axis[kBinLevels]; };
Geometry makeGeometry(const at_box &box, unsigned maxDepth)
{
    if (maxDepth > kBinLevels) throw Error(AT_EINVAL, "Binary maxDepth exceeds key precision");
    long double len[3] = {static_cast<long double>(box.xmax) - box.xmin,
                          static_cast<long double>(box.ymax) - box.ymin,
                          static_cast<long double>(box.zmax) - box.zmin};
    for (auto s : len)
        if (!(s > 0) || !std::isfinite(s)) throw Error(AT_EINVAL, "Tree box must have finite positive sides");
    Geometry g;
    for (unsigned d = 0; d < kBinLevels; ++d)
    {
        unsigned a = 0;
        if (len[1] > len[a]) a = 1;
        if (len[2] > len[a]) a = 2;
        g.axis[d] = std::uint8_t(a);
        len[a] *= 0.5L;
    }
    return g;
}
// box stored in Real and widened to double, as computeSfcKeysBinaryKernel does
// (kdtree3d_rebalance.cuh:148-160).
template <class T>
Key binaryKey(T x, T y, T z, const at_box &box, const Geometry &g)
{
    if (isEscaped(x, y, z)) return kInvalidKey;
    const double v[3] = {
        (double(x) - double(T(box.xmin))) / (double(T(box.xmax)) - double(T(box.xmin))),
        (double(y) - double(T(box.ymin))) / (double(T(box.ymax)) - double(T(box.ymin))),
        (double(z) - double(T(box.zmin))) / (double(T(box.zmax)) - double(T(box.zmin)))};
    Key f[3];
    for (int a = 0; a < 3; ++a)
    {
        const double c1 = (1.0 < v[a]) ? 1.0 : v[a];   // cmin(1, v)
        const double c0 = (0.0 > c1) ? 0.0 : c1;       // cmax(0, .)
        const Key q = Key(c0 * double(kMaxKey));
        f[a] = q < kMaxKey - 1 ? q : kMaxKey - 1;
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

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (u64 key, u32 value) pairs, OpenMP. Equal keys keep
// ascending input order, which is the contract of the reference's
// cub::DeviceRadixSort::SortPairs (octree.cu:301-304).
// ---------------------------------------------------------------------------
void sortPairs(Key *keys, std::uint32_t *vals, std::size_t n)
{
    if (n < 2) return;
    std::vector<Key> tk(n);
    std::vector<std::uint32_t> tv(n);
    Key *ks = keys, *kd = tk.data();
    std::uint32_t *vs = vals, *vd = tv.data();
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    std::vector<std::size_t> hist(std::size_t(nt) * 256);
    for (unsigned shift = 0; shift < 64; shift += 8)
    {
        std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(nt)
        {
            int t = 0;
#ifdef _OPENMP
            t = omp_get_thread_num();
#endif
            const std::size_t b = n * t / nt, e = n * (t + 1) / nt;
            std::size_t *h = &hist[std::size_t(t) * 256];
            for (std::size_t i = b; i < e; ++i) ++h[(ks[i] >> shift) & 255u];
        }
        // a pass whose digit is constant is the identity: skip it
        bool trivial = false;
        for (unsigned d = 0; d < 256 && !trivial; ++d)
        {
            std::size_t s = 0;
            for (int t = 0; t < nt; ++t) s += hist[std::size_t(t) * 256 + d];
            if (s == n) trivial = true;
            else if (s) break;
        }
        if (trivial) continue;
        std::size_t run = 0;
        for (unsigned d = 0; d < 256; ++d)
            for (int t = 0; t < nt; ++t)
            {
                const std::size_t c = hist[std::size_t(t) * 256 + d];
                hist[std::size_t(t) * 256 + d] = run;
                run += c;
            }
#pragma omp parallel num_threads(nt)
        {
            int t = 0;
#ifdef _OPENMP
            t = omp_get_thread_num();
#endif
            const std::size_t b = n * t / nt, e = n * (t + 1) / nt;
            std::size_t *o = &hist[std::size_t(t) * 256];
            for (std::size_t i = b; i < e; ++i)
            {
                const std::size_t p = o[(ks[i] >> shift) & 255u]++;
                kd[p] = ks[i];
                vd[p] = vs[i];
            }
        }
        std::swap(ks, kd);
        std::swap(vs, vd);
    }
    if (ks != keys)
    {
        std::memcpy(keys, ks, n * sizeof(Key));
        std::memcpy(vals, vs, n * sizeof(std::uint32_t));
    }
}

// std::lower_bound / upper_bound (helpers.cuh:66-107)
inline std::size_t lowerBound(const Key *a, std::size_t n, Key v)
{
    return std::size_t(std::lower_bound(a, a + n, v) - a);
}
inline std::size_t upperBound(const Key *a, std::size_t n, Key v)
{
    return std::size_t(std::upper_bound(a, a + n, v) - a);
}

// ---------------------------------------------------------------------------
// A thrust::device_vector stand-in that tracks the logical size and the
// allocation capacity; the octree's leafCapacity depends on these capacities
// (octree.cu:453-457). Thrust is a dependency of the reference that is not
// vendored: CCCL as shipped with CUDA 13.3, the reference's toolchain
// (benchmarks/env/environment.txt:4; CMakeLists.txt:15 finds Thrust). Its
// published vector_base algorithm (thrust/detail/vector_base.inl):
//   reserve(n), n > capacity: allocates exactly n elements;
//   growing to n elements without room (append / fill_insert, the path of
//   resize(n) and resize(n, no_init)): new capacity = size + max(size, n -
//   size), and at least 2 * capacity, i.e. max(2 size, n, 2 capacity);
//   shrinking or growing within the capacity keeps the allocation.
// The reference's call sites: resizeWithHeadroom (helpers.cuh:122-141),
// prepareGraphBuffers (octree.cu:182-241) and the build loop's temporaries
// (octree.cu:159-180).
// ---------------------------------------------------------------------------
template <class T>
struct TVec
{
    std::vector<T> mem;
    std::size_t sz = 0, cap = 0;
    void ensure(std::size_t n)
    {
        if (mem.size() < n) mem.resize(n);
    }
    void reserve(std::size_t n)
    {
        if (n > cap) { cap = n; ensure(cap); }
    }
    void resize(std::size_t n)
    {
        if (n > cap)
        {
            std::size_t c = std::max(2 * sz, n);
            c = std::max(c, 2 * cap);
            cap = c;
            ensure(cap);
        }
        sz = n;
        ensure(sz);
    }
    // resizeWithHeadroom (helpers.cuh:122-141)
    void resizeHeadroom(std::size_t req)
    {
        if (req > cap) reserve(req + req / 4);
        resize(req);
    }
    T *data() { ensure(1); return mem.data(); }
    const T *data() const { return mem.data(); }
    T &operator[](std::size_t i) { return mem[i]; }
    const T &operator[](std::size_t i) const { return mem[i]; }
    bool empty() const { return sz == 0; }
    void swap(TVec &o) { mem.swap(o.mem); std::swap(sz, o.sz); std::swap(cap, o.cap); }
};

// RebalanceState (helpers.cuh:110-120)
struct State
{
    int changed = 0, converged = 0, numLeaves = 1, newNumLeaves = 1, needsResize = 0,
        numNodes = 0, activeBuffer = 0;
    unsigned numActiveParticles = 0;
};

struct Views { bool structure = false, counts = false, mapping = false; };
inline bool any(const Views &v) { return v.structure || v.counts || v.mapping; }

// ---------------------------------------------------------------------------
// TreeView<b> (tree_view.cu), state mode: sizes come from the RebalanceState.
// ---------------------------------------------------------------------------
template <unsigned B>
struct TreeView
{
    static constexpr unsigned kMax = kBinLevels / B;        // 21 or 63
    static constexpr unsigned kChildren = 1u << B;
    std::size_t leafCap = 0, particleCap = 0;
    Views prepared;
    bool scratch = false;
    bool depthAllocated = false;
    std::uint8_t maxDepth = 0;
    std::size_t numNodes = 0;
    TVec<Key> sfcBoxIndex;
    TVec<std::uint8_t> hasBoxSplit, boxDepth;
    TVec<std::int32_t> parentIndex, childIndex, levelOffset, leafToNode;
    TVec<std::uint32_t> particleBegin, particleCounts;
    TVec<Key> sfcParticleIndex;

    static Key range(unsigned level) { return kMaxKey >> (B * level); }
    static unsigned levelOf(Key r) { return unsigned(clz64(r)) / B; }
    static Key startAt(Key key, unsigned level) { const Key r = range(level); return (key / r) * r; }
    static unsigned lca(Key a, Key b) // lcaLevel (tree_view.cu:94-108)
    {
        unsigned res = 0;
        for (unsigned l = 1; l <= kMax; ++l)
        {
            if (startAt(a, l) != startAt(b, l)) break;
            res = l;
        }
        return res;
    }
    static Key reversedPath(Key start, unsigned level) // childCentricLayoutKey (:120-134)
    {
        Key r = 0;
        for (unsigned d = level; d > 0; --d)
        {
            const Key ps = startAt(start, d - 1);
            const unsigned digit = unsigned((start - ps) / range(d)) % kChildren;
            r = r * Key(kChildren) + Key(digit);
        }
        return r;
    }

    // prepare (tree_view.cu:755-793)
    void prepare(std::size_t leafCapacity, std::size_t numParticles, Views cfg)
    {
        if (!any(cfg)) return;
        leafCap = std::max(leafCap, leafCapacity);
        particleCap = std::max(particleCap, numParticles);
        prepared.structure |= cfg.structure;
        prepared.counts |= cfg.counts;
        prepared.mapping |= cfg.mapping;
        const std::size_t internals = (leafCap - 1) / (kChildren - 1);
        const std::size_t nodes = leafCap + internals;
        scratch = true;
        depthAllocated = true;
        sfcBoxIndex.resizeHeadroom(nodes);
        hasBoxSplit.resizeHeadroom(nodes);
        boxDepth.resizeHeadroom(nodes);
        parentIndex.resizeHeadroom(nodes);
        childIndex.resizeHeadroom(nodes);
        leafToNode.resizeHeadroom(leafCap);
        levelOffset.resizeHeadroom(kMax + 2);
        if (prepared.counts)
        {
            particleBegin.resizeHeadroom(nodes);
            particleCounts.resizeHeadroom(nodes + 1);
        }
        if (prepared.mapping) sfcParticleIndex.resizeHeadroom(particleCap);
    }
    // syncSizes (tree_view.cu:795-812)
    void syncSizes(std::size_t numLeaves, std::size_t numActive)
    {
        if (!scratch || numLeaves > leafCap) return;
        numNodes = numLeaves + (numLeaves - 1) / (kChildren - 1);
        sfcBoxIndex.resize(numNodes);
        hasBoxSplit.resize(numNodes);
        boxDepth.resize(numNodes);
        parentIndex.resize(numNodes);
        childIndex.resize(numNodes);
        leafToNode.resize(numLeaves);
        if (prepared.counts)
        {
            particleBegin.resize(numNodes);
            particleCounts.resize(numNodes + 1);
        }
        if (prepared.mapping) sfcParticleIndex.resize(numActive);
    }
    // build (tree_view.cu:814-1134) in state mode.
    void build(const Key *K, const Key *P, const std::uint32_t *N, std::size_t numParticles,
               Views cfg, State &st)
    {
        if (!any(cfg)) return;
        if (!scratch || (cfg.counts && !prepared.counts) ||
            (cfg.mapping && (!prepared.mapping || numParticles > particleCap)))
        {
            st.needsResize = 1; // flagViewCapacityKernel (:55-58)
            return;
        }
        const int capacity = int(leafCap);
        const int L = st.numLeaves <= capacity ? st.numLeaves : 0;
        // ---- stage 1: classify (:151-203)
        if (st.numLeaves > capacity) st.needsResize = 1;
        else st.numNodes = L + (L - 1) / int(kChildren - 1);
        std::vector<std::uint8_t> leafLevel(std::size_t(std::max(L, 0)));
        std::vector<std::int32_t> leavesPerLevel(kMax + 1, 0);
        struct Cand { Key start; std::uint8_t level; };
        std::vector<Cand> cands;
        for (int i = 0; i < L; ++i)
        {
            const unsigned lv = levelOf(K[i + 1] - K[i]);
            leafLevel[std::size_t(i)] = std::uint8_t(lv);
            ++leavesPerLevel[lv];
        }
        for (int i = 0; i + 1 < L; ++i)
        {
            const unsigned lv = lca(K[i], K[i + 1]);
            const Key ps = startAt(K[i + 1], lv);
            const unsigned slot = unsigned((K[i + 1] - ps) / range(lv + 1));
            if (slot == 1) cands.push_back({ps, std::uint8_t(lv)});
        }
        if (L == 0) return;
        // ---- stage 2: level metadata (:211-275)
        unsigned maxLevel = 0;
        for (unsigned l = 0; l <= kMax; ++l)
            if (leavesPerLevel[l] != 0) maxLevel = l;
        std::vector<std::int32_t> internalPer(kMax + 1, 0), nodesPer(kMax + 1, 0),
            offInternal(kMax + 2, 0);
        maxDepth = std::uint8_t(maxLevel);
        nodesPer[maxLevel] = leavesPerLevel[maxLevel];
        for (int l = int(maxLevel) - 1; l >= 0; --l)
        {
            internalPer[l] = nodesPer[l + 1] / int(kChildren);
            nodesPer[l] = leavesPerLevel[l] + internalPer[l];
        }
        std::int32_t *lo = levelOffset.data();
        lo[0] = 0;
        for (unsigned l = 0; l <= maxLevel; ++l) lo[l + 1] = lo[l] + nodesPer[l];
        const std::int32_t total = lo[maxLevel + 1];
        for (unsigned l = maxLevel + 1; l <= kMax; ++l) lo[l + 1] = total;
        for (unsigned l = 0; l <= maxLevel; ++l) offInternal[l + 1] = offInternal[l] + internalPer[l];
        for (unsigned l = maxLevel + 1; l <= kMax; ++l) offInternal[l + 1] = offInternal[maxLevel + 1];
        // ---- stages 3-5: internal nodes grouped by level (stable), then the
        //      child-centric layout order (reversed child path) (:279-358, :651-689)
        const int activeI = (st.numLeaves - 1) / int(kChildren - 1);
        std::vector<Cand> bySfc(cands.begin(), cands.begin() + std::min<std::size_t>(cands.size(), std::size_t(std::max(activeI, 0))));
        std::stable_sort(bySfc.begin(), bySfc.end(), [](const Cand &a, const Cand &b) { return a.level < b.level; });
        const int I = int(bySfc.size());
        const std::size_t nI = bySfc.size();
        std::vector<std::pair<Key, int>> layout(nI);
        for (int i = 0; i < I; ++i)
            layout[std::size_t(i)] = {(Key(1) << (B * bySfc[std::size_t(i)].level)) |
                                          reversedPath(bySfc[std::size_t(i)].start, bySfc[std::size_t(i)].level),
                                      i};
        std::stable_sort(layout.begin(), layout.end(),
                         [](const std::pair<Key, int> &a, const std::pair<Key, int> &b) { return a.first < b.first; });
        std::vector<Key> startSfc(nI), startLayout(nI);
        std::vector<std::uint8_t> levelLayout(nI);
        std::vector<std::int32_t> mBySfc(nI);
        for (int i = 0; i < I; ++i) startSfc[std::size_t(i)] = bySfc[std::size_t(i)].start;
        for (int li = 0; li < I; ++li)
        {
            const int sp = layout[std::size_t(li)].second;
            const unsigned lv = bySfc[std::size_t(sp)].level;
            levelLayout[std::size_t(li)] = std::uint8_t(lv);
            startLayout[std::size_t(li)] = bySfc[std::size_t(sp)].start;
            mBySfc[std::size_t(sp)] = li - offInternal[lv];
        }
        // finalNodeIndex (:362-401)
        auto nodeIndex = [&](unsigned level, Key start) -> std::int32_t {
            if (level == 0) return 0;
            const Key ps = startAt(start, level - 1);
            const std::int32_t b0 = offInternal[level - 1], e0 = offInternal[level];
            const std::int32_t pos = b0 + std::int32_t(std::lower_bound(startSfc.begin() + b0, startSfc.begin() + e0, ps) - (startSfc.begin() + b0));
            const std::int32_t pm = mBySfc[std::size_t(pos)];
            const unsigned slot = unsigned((start - ps) / range(level));
            return lo[level] + std::int32_t(slot) * internalPer[level - 1] + pm;
        };
        const std::int32_t nn = st.numNodes;
        // ---- stage 6: internal nodes (:407-482)
        for (int idx = 0; idx < I; ++idx)
        {
            const unsigned lv = levelLayout[std::size_t(idx)];
            const Key start = startLayout[std::size_t(idx)];
            const Key end = start + range(lv);
            const std::int32_t ni = nodeIndex(lv, start);
            sfcBoxIndex[std::size_t(ni)] = placeholder(start, int(B * lv));
            boxDepth[std::size_t(ni)] = std::uint8_t(lv);
            hasBoxSplit[std::size_t(ni)] = 1;
            if (lv == 0) parentIndex[std::size_t(ni)] = -1;
            const std::int32_t m = idx - offInternal[lv];
            const std::int32_t M = internalPer[lv];
            childIndex[std::size_t(ni)] = lo[lv + 1] + m;
            for (unsigned s = 0; s < kChildren; ++s)
                parentIndex[std::size_t(lo[lv + 1] + std::int32_t(s) * M + m)] = ni;
            if (cfg.counts)
            {
                const std::size_t b = lowerBound(P, numParticles, start), e = lowerBound(P, numParticles, end);
                particleBegin[std::size_t(ni)] = unsigned(b);
                particleCounts[std::size_t(ni)] = unsigned(e - b);
                if (lv == 0) particleCounts[std::size_t(nn)] = unsigned(e - b);
            }
        }
        // ---- stage 7: leaves (:488-538); leaf begins = exclusive scan of N (:691-703)
        std::uint32_t running = 0;
        for (int i = 0; i < L; ++i)
        {
            const unsigned lv = leafLevel[std::size_t(i)];
            const std::int32_t ni = nodeIndex(lv, K[i]);
            sfcBoxIndex[std::size_t(ni)] = placeholder(K[i], int(B * lv));
            boxDepth[std::size_t(ni)] = std::uint8_t(lv);
            hasBoxSplit[std::size_t(ni)] = 0;
            childIndex[std::size_t(ni)] = -1;
            leafToNode[std::size_t(i)] = ni;
            if (lv == 0) parentIndex[std::size_t(ni)] = -1;
            if (cfg.counts)
            {
                particleBegin[std::size_t(ni)] = running;
                particleCounts[std::size_t(ni)] = N[i];
                if (lv == 0) particleCounts[std::size_t(nn)] = N[i];
            }
            running += N[i];
        }
        // ---- particle mapping (:542-562), every slot of the backing buffer
        if (cfg.mapping)
        {
            for (std::size_t p = 0; p < numParticles; ++p)
            {
                if (P[p] >= kMaxKey) { sfcParticleIndex[p] = kInvalidKey; continue; }
                const std::size_t leaf = upperBound(K, std::size_t(L), P[p]) - 1;
                sfcParticleIndex[p] = Key(std::int64_t(leafToNode[leaf]));
            }
        }
    }
};

// ---------------------------------------------------------------------------
// Base tree
// ---------------------------------------------------------------------------
struct TreeBase
{
    virtual ~TreeBase() = default;
    virtual void build() = 0;
    virtual void update() = 0;
    virtual void computeViews(const Views *v) = 0;
    virtual bool needsResize() = 0;
    virtual void resizeBuffers() = 0;
    virtual void buffer(int id, const void **d, std::size_t *s, std::size_t *c) = 0;
    virtual void scalars(at_scalars *o) = 0;
    virtual unsigned maxAchievedDepth() = 0;
    State st;
    std::int32_t devLeaves = 0, devNodes = 0;
    std::uint32_t devActive = 0;
};

// ---------------------------------------------------------------------------
// Octree (octree.cu + octree_rebalance.cuh)
// ---------------------------------------------------------------------------
template <class Real>
struct Octree : TreeBase
{
    // config (OctreeConfig, config.hpp:118-130)
    unsigned maxDepth = 21, bucket = 64, updateIterations = 2;
    bool useNF = false, hilb = false;
    at_box box{0, 1, 0, 1, 0, 1};
    Views views;
    Views preparedViews;
    const Real *x = nullptr, *y = nullptr, *z = nullptr;
    std::size_t n = 0, numActive = 0, numNodes = 0, leafCapacity = 0;
    bool built = false;
    int activeBuffer = 0;
    std::vector<Key> P, Psort;
    std::vector<std::uint32_t> Perm, PermSort;
    TVec<Key> K, tmpTree;
    TVec<std::uint32_t> N, NF, tmpCounts, tmpNF;
    TVec<std::uint8_t> nfDirty, groupCanMerge, tmpGroupCanMerge, groupDirty;
    TreeView<3> view;

    Key *treeOf(int buf) { return buf ? tmpTree.data() : K.data(); }

    void computeParticleKeys() // octree.cu:285-307
    {
        const FBox<Real> fb = makeBox<Real>(box);
#pragma omp parallel for schedule(static)
        for (std::size_t i = 0; i < n; ++i)
        {
            Psort[i] = isEscaped(x[i], y[i], z[i]) ? kInvalidKey : sfc3D<Real>(x[i], y[i], z[i], fb, hilb);
            PermSort[i] = std::uint32_t(i);
        }
        std::copy(Psort.begin(), Psort.end(), P.begin());
        std::copy(PermSort.begin(), PermSort.end(), Perm.begin());
        sortPairs(P.data(), Perm.data(), n);
    }
    std::size_t activeCount() const { return n ? lowerBound(P.data(), n, nodeRange(0)) : 0; }

    // siblingIndexAndLevel (octree_rebalance.cuh:26-50)
    static int siblingIndexAndLevel(const Key *tree, int numNodes, int i, unsigned &levelOut)
    {
        const Key thisNode = tree[i];
        const unsigned level = treeLevel(tree[i + 1] - thisNode);
        levelOut = level;
        if (level == 0) return -1;
        const int s = int(octalDigit(thisNode, level));
        const int first = i - s;
        if (first < 0 || first > numNodes - 8) return -1;
        const bool have = tree[first + 8] == tree[first] + nodeRange(level - 1);
        return have ? s : -1;
    }
    // calculateNodeOp (:174-249)
    int nodeOp(const Key *tree, int numNodes, int i, const std::uint32_t *counts,
               const std::uint8_t *gcm, const std::uint32_t *nf) const
    {
        unsigned level = 0;
        const int s = siblingIndexAndLevel(tree, numNodes, i, level);
        if (!useNF)
        {
            if (s > 0)
            {
                const std::uint32_t *g = counts + i - s;
                std::size_t parent = 0;
                for (int k = 0; k < 8; ++k) parent += g[k];
                if (parent <= std::size_t(bucket)) return 0;
            }
        }
        else if (s >= 0)
        {
            const int first = i - s;
            if (first + 8 <= numNodes && gcm[first]) return s == 0 ? 1 : 0;
        }
        const unsigned c = useNF ? nf[i] : counts[i];
        if (!useNF)
        {
            // unsigned products wrap like the reference's 32-bit arithmetic
            if (c > bucket * 512u && level + 3 < kOctLevels) return 4096;
            if (c > bucket * 64u && level + 2 < kOctLevels) return 512;
            if (c > bucket * 8u && level + 1 < kOctLevels) return 64;
            if (c > bucket && level < kOctLevels) return 8;
        }
        else
        {
            const std::uint64_t b = bucket;
            if (std::uint64_t(c) > b * 262144u && level + 3 < kOctLevels) return 4096;
            if (std::uint64_t(c) > b * 4096u && level + 2 < kOctLevels) return 512;
            if (std::uint64_t(c) > b * 64u && level + 1 < kOctLevels) return 64;
            if (std::uint64_t(c) > b && level < kOctLevels) return 8;
        }
        return 1;
    }
    // processNode (:253-365)
    template <class Off>
    void processNode(int i, int numOld, const Key *oldTree, const Off *ops, Key *newTree,
                     const std::uint32_t *oldCounts, const std::uint32_t *oldNF,
                     std::uint32_t *newCounts, std::uint32_t *newNF, std::uint8_t *dirty,
                     const std::uint8_t *oldGcm, std::uint8_t *newGcm, std::uint8_t *gDirty) const
    {
        auto op = [&](int j) { return int(ops[j + 1] - ops[j]); };
        const Key node = oldTree[i];
        const unsigned level = treeLevel(oldTree[i + 1] - node);
        const int code = op(i);
        const std::int64_t out = std::int64_t(ops[i]);
        if (code == 0) return;
        if (code == 1)
        {
            newTree[out] = node;
            if (useNF)
            {
                const bool leader = i + 1 < numOld && op(i + 1) == 0;
                if (leader)
                {
                    dirty[out] = 1;
                    gDirty[out] = 1;
                }
                else
                {
                    newCounts[out] = oldCounts[i];
                    newNF[out] = oldNF[i];
                    dirty[out] = 0;
                    bool reusable = false;
                    unsigned sl = 0;
                    const int s = siblingIndexAndLevel(oldTree, numOld, i, sl);
                    if (s == 0 && i + 8 <= numOld)
                    {
                        reusable = true;
                        for (int k = 0; k < 8; ++k)
                            if (op(i + k) != 1) { reusable = false; break; }
                    }
                    if (reusable) { newGcm[out] = oldGcm[i]; gDirty[out] = 0; }
                    else gDirty[out] = 1;
                }
            }
            return;
        }
        const unsigned diff = code == 8 ? 1u : log8ceil32(std::uint32_t(code));
        const Key nr = nodeRange(level + diff);
        for (int s = 0; s < code; ++s)
        {
            newTree[out + s] = node + Key(s) * nr;
            if (useNF) { dirty[out + s] = 1; gDirty[out + s] = 1; }
        }
    }
    // computeNFCooperative (:579-634): own population times the population of
    // the 27 same-level cells (periodic), counted over every slot of P.
    std::uint64_t nearField(Key start, unsigned level, std::uint32_t *own) const
    {
        unsigned c[3];
        sfcDecode(start, hilb, c);
        const int len = int(1u << (kOctLevels - level));
        const Key r = nodeRange(level);
        std::uint64_t sum = 0, center = 0;
        const int R = int(1u << kOctLevels);
        auto pbc = [R](int v) { int t = v < 0 ? v + R : v; return t >= R ? t - R : t; };
        for (int cell = 0; cell < 27; ++cell)
        {
            const int dx = cell % 3 - 1, dy = (cell / 3) % 3 - 1, dz = cell / 9 - 1;
            Key ck = start;
            if (cell != 13)
                ck = enclosingBoxCode(sfcEncode(unsigned(pbc(int(c[0]) + dx * len)),
                                                unsigned(pbc(int(c[1]) + dy * len)),
                                                unsigned(pbc(int(c[2]) + dz * len)), hilb),
                                      level);
            const std::size_t f = lowerBound(P.data(), n, ck), l = lowerBound(P.data(), n, ck + r);
            const std::uint64_t pop = std::uint64_t(l - f);
            sum += pop;
            if (cell == 13) center = pop;
        }
        if (own) *own = unsigned(std::min<std::uint64_t>(center, 0xffffffffu));
        return center * sum;
    }
    static unsigned nodeCount(Key a, Key b, const Key *P, std::size_t n) // calculateNodeCount (:368-380)
    {
        const std::size_t s = lowerBound(P, n, a), e = lowerBound(P, n, b);
        return unsigned(std::min<std::size_t>(e - s, 0xffffffffu));
    }
    // IsDirtyGroupLeader (:82-105)
    bool dirtyLeader(const Key *tree, const std::uint8_t *gDirty, int numNodes, int i) const
    {
        if (i + 8 > numNodes) return false;
        if (gDirty && !gDirty[i]) return false;
        unsigned sl = 0;
        const int s = siblingIndexAndLevel(tree, numNodes, i, sl);
        return s == 0 && sl > 0;
    }
    // NF and groupCanMerge refresh for dirty (or all) leaves and groups
    void refreshNF(const Key *tree, int L, std::uint32_t *counts, std::uint32_t *nf,
                   std::uint8_t *gcm, const std::uint8_t *dirty, const std::uint8_t *gDirty)
    {
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < L; ++i)
        {
            if (dirty && !dirty[i]) continue;
            const unsigned lv = treeLevel(tree[i + 1] - tree[i]);
            const std::uint64_t v = nearField(tree[i], lv, counts + i);
            nf[i] = unsigned(std::min<std::uint64_t>(v, 0xffffffffu));
        }
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < L; ++i)
        {
            if (!dirtyLeader(tree, gDirty, L, i)) continue;
            const unsigned lv = treeLevel(tree[i + 1] - tree[i]);
            const std::uint64_t parent = nearField(tree[i], lv - 1, nullptr);
            gcm[i] = parent <= std::uint64_t(bucket) ? 1 : 0;
        }
    }
    void refreshCounts(const Key *tree, int L, std::uint32_t *counts)
    {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < L; ++i) counts[i] = nodeCount(tree[i], tree[i + 1], P.data(), n);
    }

    // one build pass: updateOctreeGpu (:871-985); returns true when converged
    bool buildPass()
    {
        const int L = int(K.sz - 1);
        const Key *tree = K.data();
        std::vector<std::int32_t> ops(std::size_t(L) + 1);
        st.changed = 0;
        for (int i = 0; i < L; ++i)
        {
            ops[std::size_t(i)] = nodeOp(tree, L, i, N.data(), groupCanMerge.data(), NF.data());
            if (ops[std::size_t(i)] != 1) st.changed = 1;
        }
        ops[std::size_t(L)] = 0;
        if (!st.converged && st.changed == 0) st.converged = 1;
        std::int32_t run = 0;
        for (int i = 0; i <= L; ++i) { const std::int32_t v = ops[std::size_t(i)]; ops[std::size_t(i)] = run; run += v; }
        if (!st.changed) return true;
        const int newL = ops[std::size_t(L)];
        tmpTree.resizeHeadroom(std::size_t(newL) + 1);
        if (useNF)
        {
            tmpCounts.resizeHeadroom(std::size_t(newL));
            tmpNF.resizeHeadroom(std::size_t(newL));
            nfDirty.resizeHeadroom(std::size_t(newL));
            tmpGroupCanMerge.resizeHeadroom(std::size_t(newL));
            groupDirty.resizeHeadroom(std::size_t(newL));
        }
        for (int i = 0; i < L; ++i)
            processNode(i, L, tree, ops.data(), tmpTree.data(), N.data(), NF.data(), tmpCounts.data(),
                        tmpNF.data(), nfDirty.data(), groupCanMerge.data(), tmpGroupCanMerge.data(),
                        groupDirty.data());
        tmpTree[std::size_t(newL)] = nodeRange(0);
        K.swap(tmpTree);
        if (useNF)
        {
            N.swap(tmpCounts);
            NF.swap(tmpNF);
            groupCanMerge.swap(tmpGroupCanMerge);
            refreshNF(K.data(), newL, N.data(), NF.data(), groupCanMerge.data(), nfDirty.data(), groupDirty.data());
        }
        else
        {
            N.resizeHeadroom(std::size_t(newL));
            refreshCounts(K.data(), newL, N.data());
        }
        return false;
    }

    // prepareGraphBuffers (octree.cu:182-241)
    void prepareGraphBuffers(std::size_t capacity)
    {
        constexpr std::size_t maxLeaves = (std::size_t(std::numeric_limits<int>::max()) - 1) / 8 * 7 + 1;
        if (capacity > maxLeaves) throw Error(AT_ELENGTH, "adaptive_octree: buffer capacity exceeds node index range");
        const std::size_t leaves = K.sz - 1;
        leafCapacity = capacity;
        auto prep = [](auto &v, std::size_t s) { if (v.cap < s) v.reserve(s); v.resize(s); };
        prep(K, capacity + 1); prep(tmpTree, capacity + 1);
        prep(N, capacity); prep(NF, capacity); prep(tmpCounts, capacity); prep(tmpNF, capacity);
        prep(nfDirty, capacity); prep(groupCanMerge, capacity); prep(tmpGroupCanMerge, capacity);
        prep(groupDirty, capacity);
        view.prepare(capacity, n, preparedViews);
        K.resize(leaves + 1); N.resize(leaves); NF.resize(leaves);
        tmpTree.resize(leaves + 1); tmpCounts.resize(leaves); tmpNF.resize(leaves);
        view.syncSizes(leaves, numActive);
    }

    // graph-capturable pass: updateOctreeGraphGpu (:1275-1322)
    void graphRefresh(bool full)
    {
        if (st.converged) return;
        const int L = st.numLeaves;
        const int ab = st.activeBuffer;
        Key *tree = treeOf(ab);
        std::uint32_t *counts = ab ? tmpCounts.data() : N.data();
        if (!useNF) { refreshCounts(tree, L, counts); return; }
        std::uint32_t *nf = ab ? tmpNF.data() : NF.data();
        std::uint8_t *gcm = ab ? tmpGroupCanMerge.data() : groupCanMerge.data();
        refreshNF(tree, L, counts, nf, gcm, full ? nullptr : nfDirty.data(), full ? nullptr : groupDirty.data());
    }
    void graphPass()
    {
        const int cap = int(leafCapacity);
        st.changed = 0;
        std::vector<std::int64_t> ops(std::size_t(cap) + 1, 0);
        const int ab = st.activeBuffer;
        const Key *tree = treeOf(ab);
        const std::uint32_t *counts = ab ? tmpCounts.data() : N.data();
        const std::uint32_t *nf = ab ? tmpNF.data() : NF.data();
        const std::uint8_t *gcm = ab ? tmpGroupCanMerge.data() : groupCanMerge.data();
        if (!st.converged && !st.needsResize)
            for (int i = 0; i < st.numLeaves; ++i)
            {
                const int d = nodeOp(tree, st.numLeaves, i, counts, gcm, nf);
                if (d != 1) st.changed = 1;
                ops[std::size_t(i)] = d;
            }
        std::int64_t run = 0;
        for (auto &v : ops) { const std::int64_t t = v; v = run; run += t; }
        // graphCheckCapacityKernel (:1211-1232)
        if (!st.converged)
        {
            if (st.needsResize || !st.changed) st.converged = 1;
            else if (ops[std::size_t(cap)] > cap) { st.needsResize = 1; st.converged = 1; }
            else st.newNumLeaves = int(ops[std::size_t(cap)]);
        }
        if (st.converged) { graphRefresh(false); return; }
        // graphProcessNodesKernel (:1234-1265) into the inactive buffer
        Key *out = treeOf(ab ^ 1);
        std::uint32_t *tc = ab ? N.data() : tmpCounts.data();
        std::uint32_t *tnf = ab ? NF.data() : tmpNF.data();
        std::uint8_t *tg = ab ? groupCanMerge.data() : tmpGroupCanMerge.data();
        for (int i = 0; i < st.numLeaves; ++i)
            processNode(i, st.numLeaves, tree, ops.data(), out, counts, nf, tc, tnf, nfDirty.data(), gcm, tg,
                        groupDirty.data());
        out[st.newNumLeaves] = nodeRange(0);
        // graphCommitNodesKernel (:1267-1273)
        st.numLeaves = st.newNumLeaves;
        st.activeBuffer ^= 1;
        graphRefresh(false);
    }

    void buildTreeView(Views v) // octree.cu:309-318
    {
        preparedViews.structure |= v.structure;
        preparedViews.counts |= v.counts;
        preparedViews.mapping |= v.mapping;
        const int ab = st.activeBuffer;
        view.build(treeOf(ab), P.data(), ab ? tmpCounts.data() : N.data(), n, v, st);
    }

    void build() override // octree.cu:428-469
    {
        K.resize(2); N.resize(1); NF.resize(1);
        st = State{};
        st.numLeaves = st.newNumLeaves = 0; // cudaMemsetAsync(state, 0)
        computeParticleKeys();
        K[0] = 0; K[1] = nodeRange(0);
        {
            const std::uint64_t active = activeCount();
            N[0] = unsigned(active);
            const std::uint64_t sq = active * active;
            NF[0] = sq > 0xffffffffull / 27ull ? 0xffffffffu : unsigned(27ull * sq);
        }
        st.changed = 0; st.converged = 0; st.numActiveParticles = unsigned(activeCount());
        for (unsigned i = 0; i < maxDepth; ++i)
            if (buildPass()) break;
        const std::size_t leaves = K.sz - 1;
        prepareGraphBuffers(std::max({K.cap - 1, tmpTree.empty() ? std::size_t(0) : tmpTree.cap - 1,
                                      leaves + (leaves + 3) / 4}));
        st.numLeaves = st.newNumLeaves = int(leaves);
        st.changed = st.converged = st.needsResize = 0;
        st.numNodes = 0;
        st.activeBuffer = 0;
        built = true;
        buildTreeView(views);
        (void)needsResize();
    }
    void update() override // octree.cu:472-481, :243-282
    {
        if (!built) throw Error(AT_ELOGIC, "adaptive_octree: call build() before update()");
        computeParticleKeys();
        st.changed = 0; st.converged = 0; st.numActiveParticles = unsigned(activeCount());
        graphRefresh(true);
        for (unsigned i = 0; i < updateIterations; ++i) graphPass();
        buildTreeView(views);
    }
    void computeViews(const Views *v) override
    {
        if (!built) throw Error(AT_ELOGIC, "adaptive_octree: call build() before computeViews()");
        buildTreeView(v ? *v : views);
    }
    bool needsResize() override // octree.cu:499-537
    {
        if (!built) return false;
        activeBuffer = st.activeBuffer;
        const std::size_t L = std::size_t(st.numLeaves);
        K.resize(L + 1); tmpTree.resize(L + 1);
        tmpCounts.resize(L); tmpNF.resize(L); N.resize(L); NF.resize(L);
        numActive = st.numActiveParticles;
        view.syncSizes(L, numActive);
        numNodes = std::size_t(st.numNodes);
        return st.needsResize != 0;
    }
    void resizeBuffers() override // octree.cu:540-551
    {
        if (!built) throw Error(AT_ELOGIC, "adaptive_octree: call build() before resizeBuffers()");
        (void)needsResize();
        prepareGraphBuffers(leafCapacity + (leafCapacity + 3) / 4);
        st.needsResize = 0;
    }
    void buffer(int id, const void **d, std::size_t *s, std::size_t *c) override
    {
        *d = nullptr; *s = 0; *c = 0;
        switch (id)
        {
        case AT_BUF_PARTICLE_KEYS: *d = P.data(); *s = *c = n; break;
        case AT_BUF_PERM: *d = Perm.data(); *s = *c = n; break;
        case AT_BUF_CORNERSTONE: { auto &v = activeBuffer ? tmpTree : K; *d = v.data(); *s = v.sz; *c = v.cap; } break;
        case AT_BUF_LEAF_COUNTS: { auto &v = activeBuffer ? tmpCounts : N; *d = v.data(); *s = v.sz; *c = v.cap; } break;
        case AT_BUF_NEAR_FIELD_COUNTS: { auto &v = activeBuffer ? tmpNF : NF; *d = v.data(); *s = v.sz; *c = v.cap; } break;
        default: viewBuffer(view, id, d, s, c);
        }
    }
    template <class V>
    static void viewBuffer(V &view, int id, const void **d, std::size_t *s, std::size_t *c)
    {
        auto out = [&](auto &v) { *d = v.mem.empty() ? nullptr : v.mem.data(); *s = v.sz; *c = v.cap; };
        switch (id)
        {
        case AT_BUF_SFC_BOX_INDEX: out(view.sfcBoxIndex); break;
        case AT_BUF_HAS_BOX_SPLIT: out(view.hasBoxSplit); break;
        case AT_BUF_BOX_DEPTH: out(view.boxDepth); break;
        case AT_BUF_PARENT_INDEX: out(view.parentIndex); break;
        case AT_BUF_CHILD_INDEX: out(view.childIndex); break;
        case AT_BUF_LEVEL_OFFSET: out(view.levelOffset); break;
        case AT_BUF_PARTICLE_BEGIN_INDEX: out(view.particleBegin); break;
        case AT_BUF_PARTICLE_COUNTS: out(view.particleCounts); break;
        case AT_BUF_SFC_PARTICLE_INDEX: out(view.sfcParticleIndex); break;
        default: throw Error(AT_EINVAL, "unknown buffer id");
        }
    }
    void scalars(at_scalars *o) override
    {
        o->num_particles = n;
        o->num_active_particles = numActive;
        o->num_nodes = numNodes;
        o->num_leaves = (activeBuffer ? tmpCounts.sz : N.sz);
        o->leaf_capacity = leafCapacity;
        o->tree_type = AT_TREE_OCTREE;
        o->split_criterion = useNF ? AT_NF_COUNT : AT_LEAF_COUNT;
        o->sfc_kind = hilb ? AT_HILBERT : AT_MORTON;
        o->real_bytes = int(sizeof(Real));
        o->active_buffer = activeBuffer;
        o->update_iterations = int(updateIterations);
        o->max_depth = int(maxDepth);
        o->max_particles_per_leaf = bucket;
    }
    unsigned maxAchievedDepth() override { return view.depthAllocated ? view.maxDepth : 0; }
};

// ---------------------------------------------------------------------------
// Binary tree (kdtree3d.cu + kdtree3d_rebalance.cuh + binary_node.cuh)
// ---------------------------------------------------------------------------
template <class Real>
struct Binary : TreeBase
{
    unsigned maxDepth = 63, bucket = 64, updateIterations = 6;
    at_box box{0, 1, 0, 1, 0, 1};
    Geometry geometry;
    Views views, preparedViews;
    const Real *x = nullptr, *y = nullptr, *z = nullptr;
    std::size_t n = 0, numActive = 0, numNodes = 0, leafCapacity = 0;
    bool built = false;
    int activeBuffer = 0;
    std::vector<Key> P;
    std::vector<std::uint32_t> Perm;
    TVec<Key> K, tmpTree;
    TVec<std::uint32_t> N, tmpCounts;
    TreeView<1> view;

    Key *treeOf(int b) { return b ? tmpTree.data() : K.data(); }
    std::uint32_t *countsOf(int b) { return b ? tmpCounts.data() : N.data(); }

    void computeParticleKeys()
    {
#pragma omp parallel for schedule(static)
        for (std::size_t i = 0; i < n; ++i)
        {
            P[i] = binaryKey<Real>(x[i], y[i], z[i], box, geometry);
            Perm[i] = std::uint32_t(i);
        }
        sortPairs(P.data(), Perm.data(), n);
    }
    void syncSizes(std::size_t leaves)
    {
        K.resize(leaves + 1); tmpTree.resize(leaves + 1);
        N.resize(leaves); tmpCounts.resize(leaves);
        view.syncSizes(leaves, numActive);
    }
    void prepareViews()
    {
        if (!any(preparedViews)) return;
        view.prepare(leafCapacity, n, preparedViews);
    }
    void prepareGraphBuffers(std::size_t capacity) // kdtree3d.cu:96-116
    {
        constexpr std::size_t maxLeaves = std::size_t(std::numeric_limits<int>::max()) / 2 + 1;
        if (capacity > maxLeaves) throw Error(AT_ELENGTH, "adaptive_octree: buffer capacity exceeds node index range");
        const std::size_t leaves = K.sz - 1;
        leafCapacity = capacity;
        auto prep = [](auto &v, std::size_t s) { if (v.cap < s) v.reserve(s); v.resize(s); };
        prep(K, capacity + 1); prep(tmpTree, capacity + 1);
        prep(N, capacity); prep(tmpCounts, capacity);
        if (built) prepareViews();
        syncSizes(leaves);
    }
    void readState() // kdtree3d.cu:127-137
    {
        activeBuffer = st.activeBuffer;
        numActive = st.numActiveParticles;
        numNodes = std::size_t(st.numNodes);
        syncSizes(std::size_t(st.numLeaves));
    }
    void refreshCounts() // kdtree3d.cu:139-146, kdtree3d_rebalance.cuh:13-50
    {
        st.changed = 0; st.converged = 0;
        st.numActiveParticles = n ? unsigned(lowerBound(P.data(), n, kMaxKey)) : 0;
        leafCounts();
    }
    void leafCounts()
    {
        if (st.converged) return;
        const Key *tree = treeOf(st.activeBuffer);
        std::uint32_t *c = countsOf(st.activeBuffer);
        const std::size_t na = st.numActiveParticles;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < st.numLeaves; ++i)
            c[i] = na ? unsigned(lowerBound(P.data(), na, tree[i + 1]) - lowerBound(P.data(), na, tree[i])) : 0;
    }
    static bool siblings(const Key *tree, int count, int first) // binarySiblings (binary_node.cuh:91-98)
    {
        if (first < 0 || first + 1 >= count) return false;
        const Key r = tree[first + 1] - tree[first];
        return !(tree[first] & r) && tree[first + 2] - tree[first + 1] == r;
    }
    void rebalancePass() // updateBinaryGraphGpu (kdtree3d_rebalance.cuh:127-146)
    {
        const int cap = int(leafCapacity);
        st.changed = 0;
        std::vector<std::int64_t> ops(std::size_t(cap) + 1, 0);
        const int ab = st.activeBuffer;
        const Key *tree = treeOf(ab);
        const std::uint32_t *c = countsOf(ab);
        if (!st.converged && !st.needsResize)
            for (int i = 0; i < st.numLeaves; ++i)
            {
                const bool merge = i && siblings(tree, st.numLeaves, i - 1) &&
                                   std::uint64_t(c[i - 1]) + c[i] <= bucket;
                const int op = merge ? 0 : (c[i] > bucket && unsigned(clz64(tree[i + 1] - tree[i])) < maxDepth ? 2 : 1);
                if (op != 1) st.changed = 1;
                ops[std::size_t(i)] = op;
            }
        std::int64_t run = 0;
        for (auto &v : ops) { const std::int64_t t = v; v = run; run += t; }
        if (!st.converged)
        {
            if (st.needsResize || !st.changed) st.converged = 1;
            else if (ops[std::size_t(cap)] > cap) { st.needsResize = 1; st.converged = 1; }
            else st.newNumLeaves = int(ops[std::size_t(cap)]);
        }
        if (!st.converged)
        {
            Key *out = treeOf(ab ^ 1);
            out[st.newNumLeaves] = kMaxKey;
            for (int i = 0; i < st.numLeaves; ++i)
            {
                const int op = int(ops[std::size_t(i) + 1] - ops[std::size_t(i)]);
                if (!op) continue;
                const std::int64_t t = ops[std::size_t(i)];
                out[t] = tree[i];
                if (op == 2) out[t + 1] = tree[i] + (tree[i + 1] - tree[i]) / 2;
            }
            st.numLeaves = st.newNumLeaves;
            st.numNodes = (st.numLeaves - 1) * 2 + 1;
            st.activeBuffer ^= 1;
        }
        leafCounts();
    }
    void growBuffers()
    {
        prepareGraphBuffers(leafCapacity + (leafCapacity + 3) / 4);
        st.needsResize = 0;
    }
    void buildTreeView(Views v)
    {
        preparedViews.structure |= v.structure;
        preparedViews.counts |= v.counts;
        preparedViews.mapping |= v.mapping;
        const int ab = st.activeBuffer;
        view.build(treeOf(ab), P.data(), countsOf(ab), n, v, st);
    }
    void build() override // kdtree3d.cu:227-257
    {
        built = false;
        K.resize(2);
        prepareGraphBuffers(std::max<std::size_t>(2, leafCapacity));
        computeParticleKeys();
        K[0] = 0; K[1] = kMaxKey;
        st = State{};
        st.numNodes = 1;
        refreshCounts();
        for (;;)
        {
            rebalancePass();
            readState();
            if (st.needsResize) { growBuffers(); refreshCounts(); }
            else if (st.converged) break;
        }
        built = true;
        prepareViews();
        buildTreeView(views);
        if (any(views)) (void)needsResize();
    }
    void update() override
    {
        if (!built) throw Error(AT_ELOGIC, "adaptive_octree: call build() before update()");
        computeParticleKeys();
        refreshCounts();
        for (unsigned i = 0; i < updateIterations; ++i) rebalancePass();
        buildTreeView(views);
    }
    void computeViews(const Views *v) override
    {
        if (!built) throw Error(AT_ELOGIC, "adaptive_octree: call build() before computeViews()");
        buildTreeView(v ? *v : views);
    }
    bool needsResize() override
    {
        if (!built) return false;
        readState();
        return st.needsResize != 0;
    }
    void resizeBuffers() override
    {
        if (!built) throw Error(AT_ELOGIC, "adaptive_octree: call build() before resizeBuffers()");
        readState();
        growBuffers();
    }
    void buffer(int id, const void **d, std::size_t *s, std::size_t *c) override
    {
        *d = nullptr; *s = 0; *c = 0;
        switch (id)
        {
        case AT_BUF_PARTICLE_KEYS: *d = P.data(); *s = *c = n; break;
        case AT_BUF_PERM: *d = Perm.data(); *s = *c = n; break;
        case AT_BUF_CORNERSTONE: { auto &v = activeBuffer ? tmpTree : K; *d = v.data(); *s = v.sz; *c = v.cap; } break;
        case AT_BUF_LEAF_COUNTS: { auto &v = activeBuffer ? tmpCounts : N; *d = v.data(); *s = v.sz; *c = v.cap; } break;
        case AT_BUF_NEAR_FIELD_COUNTS: throw Error(AT_ELOGIC, "Tree: Binary NFCount is not implemented yet");
        default: Octree<Real>::viewBuffer(view, id, d, s, c);
        }
    }
    void scalars(at_scalars *o) override
    {
        o->num_particles = n;
        o->num_active_particles = numActive;
        o->num_nodes = numNodes;
        o->num_leaves = (activeBuffer ? tmpCounts.sz : N.sz);
        o->leaf_capacity = leafCapacity;
        o->tree_type = AT_TREE_BINARY;
        o->split_criterion = AT_LEAF_COUNT;
        o->sfc_kind = AT_MORTON;
        o->real_bytes = int(sizeof(Real));
        o->active_buffer = activeBuffer;
        o->update_iterations = int(updateIterations);
        o->max_depth = int(maxDepth);
        o->max_particles_per_leaf = bucket;
    }
    unsigned maxAchievedDepth() override // kdtree3d.cu:404-415
    {
        if (any(views)) return view.depthAllocated ? view.maxDepth : 0;
        const TVec<Key> &k = activeBuffer ? tmpTree : K;
        unsigned d = 0;
        for (std::size_t i = 0; i + 1 < k.sz; ++i) d = std::max(d, unsigned(clz64(k[i + 1] - k[i])));
        return d;
    }
};

template <class Real>
TreeBase *create(int type, const Real *x, const Real *y, const Real *z, std::size_t n, const at_config &c)
{
    // Tree::backendConfig validation (tree.hpp:108-133)
    if (c.max_depth < -1 || c.update_iterations < -1) throw Error(AT_EINVAL, "Tree: use -1 for backend defaults");
    if (n > 0 && (!x || !y || !z)) throw Error(AT_EINVAL, "adaptive_octree: coordinate device pointers must not be null");
    Views v{c.views[0] != 0, c.views[1] != 0, c.views[2] != 0};
    if (type == AT_TREE_OCTREE)
    {
        auto t = std::make_unique<Octree<Real>>();
        if (c.max_depth >= 0) t->maxDepth = unsigned(c.max_depth);
        if (c.update_iterations >= 0) t->updateIterations = unsigned(c.update_iterations);
        t->bucket = c.max_particles_per_leaf;
        t->useNF = c.split_criterion == AT_NF_COUNT;
        t->hilb = c.sfc_kind == AT_HILBERT;
        t->box = c.box;
        t->views = t->preparedViews = v;
        if (t->bucket == 0) throw Error(AT_EINVAL, "adaptive_octree: particle limit must be positive");
        if (t->updateIterations == 0) throw Error(AT_EINVAL, "adaptive_octree: updateIterations must be positive");
        t->x = x; t->y = y; t->z = z; t->n = n;
        if (n > std::size_t(std::numeric_limits<std::uint32_t>::max()))
            throw Error(AT_EINVAL, "adaptive_octree: particle count exceeds 32-bit permutation index range");
        t->P.resize(n); t->Perm.resize(n); t->Psort.resize(n); t->PermSort.resize(n);
        t->K.resize(2); t->N.resize(1); t->NF.resize(1);
        return t.release();
    }
    if (type != AT_TREE_BINARY) throw Error(AT_EINVAL, "unknown tree type");
    if (c.split_criterion != AT_LEAF_COUNT) throw Error(AT_EINVAL, "Tree: Binary NFCount is not implemented yet");
    if (c.sfc_kind != AT_MORTON) throw Error(AT_EINVAL, "Tree: Binary uses longest-side keys, not Hilbert");
    auto t = std::make_unique<Binary<Real>>();
    if (c.max_depth >= 0) t->maxDepth = unsigned(c.max_depth);
    if (c.update_iterations >= 0) t->updateIterations = unsigned(c.update_iterations);
    t->bucket = c.max_particles_per_leaf;
    t->box = c.box;
    t->views = t->preparedViews = v;
    if (n > 0 && (!x || !y || !z)) throw Error(AT_EINVAL, "adaptive_octree: coordinate device pointers must not be null");
    if (t->bucket == 0) throw Error(AT_EINVAL, "adaptive_octree: particle limit must be positive");
    if (t->updateIterations == 0) throw Error(AT_EINVAL, "adaptive_octree: updateIterations must be positive");
    t->geometry = makeGeometry(c.box, t->maxDepth);
    t->x = x; t->y = y; t->z = z; t->n = n;
    if (n > std::size_t(std::numeric_limits<std::uint32_t>::max()))
        throw Error(AT_EINVAL, "adaptive_octree: particle count exceeds 32-bit permutation index range");
    t->P.resize(n); t->Perm.resize(n);
    return t.release();
}

template <class F>
int guard(F &&f)
{
    try
    {
        f();
        g_error.clear();
        return AT_OK;
    }
    catch (const Error &e) { g_error = e.what(); return e.code; }
    catch (const std::bad_alloc &e) { g_error = e.what(); return AT_ENOMEM; }
    catch (const std::exception &e) { g_error = e.what(); return AT_EINVAL; }
}
} // namespace orc

struct orc_tree
{
    std::unique_ptr<orc::TreeBase> t;
};

using namespace orc;

extern "C" {
int orc_tree_create(int rb, int type, const void *x, const void *y, const void *z, size_t n,
                    const at_config *c, orc_tree **out)
{
    return guard([&] {
        if (!c || !out) throw Error(AT_EINVAL, "null argument");
        auto h = std::make_unique<orc_tree>();
        if (rb == 4)
            h->t.reset(create<float>(type, (const float *)x, (const float *)y, (const float *)z, n, *c));
        else if (rb == 8)
            h->t.reset(create<double>(type, (const double *)x, (const double *)y, (const double *)z, n, *c));
        else
            throw Error(AT_EINVAL, "real_bytes must be 4 or 8");
        *out = h.release();
    });
}
int orc_tree_destroy(orc_tree *t) { delete t; return AT_OK; }
int orc_tree_build(orc_tree *t) { return guard([&] { t->t->build(); }); }
int orc_tree_update(orc_tree *t) { return guard([&] { t->t->update(); }); }
int orc_tree_compute_views(orc_tree *t, const uint8_t *v)
{
    return guard([&] {
        if (v) { Views vv{v[0] != 0, v[1] != 0, v[2] != 0}; t->t->computeViews(&vv); }
        else t->t->computeViews(nullptr);
    });
}
int orc_tree_needs_resize(orc_tree *t, int *o) { return guard([&] { *o = t->t->needsResize() ? 1 : 0; }); }
int orc_tree_resize_buffers(orc_tree *t) { return guard([&] { t->t->resizeBuffers(); }); }
int orc_tree_buffer(orc_tree *t, int id, const void **d, size_t *s, size_t *c)
{
    return guard([&] { t->t->buffer(id, d, s, c); });
}
int orc_tree_scalars(orc_tree *t, at_scalars *o) { return guard([&] { t->t->scalars(o); }); }
int orc_tree_device_counters(orc_tree *t, const int32_t **l, const int32_t **nn, const uint32_t **a)
{
    return guard([&] {
        t->t->devLeaves = t->t->st.numLeaves;
        t->t->devNodes = t->t->st.numNodes;
        t->t->devActive = t->t->st.numActiveParticles;
        *l = &t->t->devLeaves; *nn = &t->t->devNodes; *a = &t->t->devActive;
    });
}
int orc_tree_max_achieved_depth(orc_tree *t, uint32_t *d) { return guard([&] { *d = t->t->maxAchievedDepth(); }); }
const char *orc_last_error(void) { return g_error.c_str(); }

int orc_sfc_keys(int rb, const void *x, const void *y, const void *z, size_t n, const at_box *box, int h,
                 uint64_t *keys)
{
    return guard([&] {
        if (rb == 4)
        {
            const FBox<float> fb = makeBox<float>(*box);
            const float *X = (const float *)x, *Y = (const float *)y, *Z = (const float *)z;
            for (size_t i = 0; i < n; ++i)
                keys[i] = isEscaped(X[i], Y[i], Z[i]) ? kInvalidKey : sfc3D<float>(X[i], Y[i], Z[i], fb, h != 0);
        }
        else
        {
            const FBox<double> fb = makeBox<double>(*box);
            const double *X = (const double *)x, *Y = (const double *)y, *Z = (const double *)z;
            for (size_t i = 0; i < n; ++i)
                keys[i] = isEscaped(X[i], Y[i], Z[i]) ? kInvalidKey : sfc3D<double>(X[i], Y[i], Z[i], fb, h != 0);
        }
    });
}
int orc_binary_keys(int rb, const void *x, const void *y, const void *z, size_t n, const at_box *box,
                    uint32_t maxDepth, uint64_t *keys)
{
    return guard([&] {
        const Geometry g = makeGeometry(*box, maxDepth);
        for (size_t i = 0; i < n; ++i)
            keys[i] = rb == 4 ? binaryKey<float>(((const float *)x)[i], ((const float *)y)[i], ((const float *)z)[i], *box, g)
                              : binaryKey<double>(((const double *)x)[i], ((const double *)y)[i], ((const double *)z)[i], *box, g);
    });
}
uint64_t orc_morton(uint32_t x, uint32_t y, uint32_t z) { return morton(x, y, z); }
uint64_t orc_hilbert(uint32_t x, uint32_t y, uint32_t z) { return hilbert(x, y, z); }
void orc_hilbert_decode(uint64_t k, uint32_t *xyz) { unsigned c[3]; hilbertDecode(k, c); xyz[0] = c[0]; xyz[1] = c[1]; xyz[2] = c[2]; }
void orc_morton_decode(uint64_t k, uint32_t *xyz) { unsigned c[3]; mortonDecode(k, c); xyz[0] = c[0]; xyz[1] = c[1]; xyz[2] = c[2]; }
int orc_sort_pairs(uint64_t *k, uint32_t *v, size_t n) { return guard([&] { sortPairs(k, v, n); }); }
uint64_t orc_near_field(const uint64_t *keys, size_t n, uint64_t start, uint32_t level, int h, uint32_t *own)
{
    Octree<double> t;
    t.P.assign(keys, keys + n);
    t.n = n;
    t.hilb = h != 0;
    return t.nearField(start, level, own);
}
int orc_binary_geometry(const at_box *box, uint32_t maxDepth, uint8_t *axis)
{
    return guard([&] {
        const Geometry g = makeGeometry(*box, maxDepth);
        std::memcpy(axis, g.axis, kBinLevels);
    });
}
int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
}
