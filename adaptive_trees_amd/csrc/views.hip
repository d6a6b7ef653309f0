// views.hip -- TreeView node arrays on gfx950 (src/common/tree_view.cu).
//
// Layout (the reference's child-centric BFS, tree_view.cu:362-401): nodes
// are grouped by level; inside a level a node's index is
//   levelOffset[l] + childSlot * I(l-1) + m(parent),
// where m ranks the internal nodes of a level by their reversed child path.
// Stages:
//   classify   leaf levels, per-level histogram, and one placeholder key per
//              internal node (the boundary between its child 0 and child 1
//              subtrees, tree_view.cu:151-203), appended in any order;
//   metadata   level offsets from the histogram (one wave, <= 64 levels);
//   sort S     internal placeholder keys: sorted = grouped by level, SFC
//              order inside a level (replaces the stable level sort);
//   layout     sort (level | reversed path) keys to get m for every node;
//   rows       internal rows, leaf rows (leaf particle begins from a scan
//              of N), then the particle -> leaf-node mapping.
// The parent of a node is found by binary search of (its placeholder key
// >> bits) in the parent level's slice of S.
#include "views.hpp"

#include "buffers.hpp"
#include "profile.hpp"

namespace at
{
namespace
{
constexpr int kT = 256;
constexpr std::uint64_t kAgg = std::uint64_t(1) << 62;
constexpr std::uint64_t kPre = std::uint64_t(2) << 62;
constexpr std::uint64_t kFlagMask = std::uint64_t(3) << 62;
constexpr std::uint64_t kSumMask = (std::uint64_t(1) << 62) - 1;

__device__ inline unsigned viewLevel(Key range, unsigned bits) { return unsigned(clz64(range)) / bits; }
__device__ inline Key viewRange(unsigned level, unsigned bits) { return kMaxKey >> (bits * level); }
__device__ inline int activeLeaves(int cap, const DevState *st) { return st->numLeaves <= cap ? st->numLeaves : 0; }
__device__ inline int activeInternals(int icap, const DevState *st, unsigned bits)
{
    const int c = (st->numLeaves - 1) / int((1u << bits) - 1);
    return c <= icap ? c : 0;
}
__device__ inline unsigned phLevel(Key ph, unsigned bits) { return unsigned(63 - clz64(ph)) / bits; }
__device__ inline Key phStart(Key ph, unsigned level, unsigned bits)
{
    const unsigned pl = bits * level;
    const Key path = ph & ((Key(1) << pl) - 1);
    return pl ? path << (63 - pl) : Key(0);
}
// reverse the order of the `level` b-bit digits of the path
__device__ inline Key reversedPath(Key ph, unsigned level, unsigned bits)
{
    Key path = ph & ((Key(1) << (bits * level)) - 1);
    Key r = 0;
    const Key mask = (Key(1) << bits) - 1;
    for (unsigned d = 0; d < level; ++d)
    {
        r = (r << bits) | (path & mask);
        path >>= bits;
    }
    return r;
}

__global__ void flagResizeKernel(DevState *st) { st->needsResize = 1; }

__global__ void prepKernel(int *leavesPerLevel, int levels, unsigned *candCount, Key *cand, std::size_t icap)
{
    const std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < std::size_t(levels)) leavesPerLevel[i] = 0;
    if (i == 0) *candCount = 0;
    for (std::size_t j = i; j < icap; j += std::size_t(gridDim.x) * blockDim.x) cand[j] = kInvalidKey;
}

// Leaf levels and their histogram, and one placeholder key per internal
// node (emitted at the boundary between its child-0 and child-1 subtrees).
// A block takes kClassifyRows rows of kT leaves: each wave adds a row's
// levels to the block histogram with one LDS atomic per distinct level, and
// the block's candidates take their places with one atomic per block (the
// list's order is free: it is sorted next).
constexpr int kClassifyRows = 8;
__global__ __launch_bounds__(kT) void classifyKernel(const Key *K0, const Key *K1, DevState *st, int cap,
                                                     unsigned bits, std::uint8_t *leafLevel, int *leavesPerLevel,
                                                     Key *cand, unsigned *candCount)
{
    __shared__ int hist[64];
    __shared__ unsigned waveCount[kClassifyRows][kT / 64], blockBase;
    if (threadIdx.x < 64) hist[threadIdx.x] = 0;
    __syncthreads();
    const Key *K = st->activeBuffer ? K1 : K0;
    const int L = activeLeaves(cap, st);
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        if (st->numLeaves > cap) st->needsResize = 1;
        else st->numNodes = L + int(unsigned(L - 1) / ((1u << bits) - 1));
    }
    const unsigned lane = __lane_id();
    const int wave = int(threadIdx.x >> 6);
    const unsigned maxLevel = 63u / bits;
    Key ph[kClassifyRows];
    std::uint32_t emitRows = 0;
#pragma unroll
    for (int r = 0; r < kClassifyRows; ++r)
    {
        const int i = int(blockIdx.x) * kT * kClassifyRows + r * kT + int(threadIdx.x);
        int lvl = -1;
        if (i < L)
        {
            lvl = int(viewLevel(K[i + 1] - K[i], bits));
            leafLevel[i] = std::uint8_t(lvl);
        }
        std::uint64_t pending = __ballot(lvl >= 0);
        while (pending)
        {
            const int leader = __ffsll((long long)pending) - 1;
            const int l0 = __shfl(lvl, leader, 64);
            const std::uint64_t same = __ballot(lvl == l0) & pending;
            if (int(lane) == leader) atomicAdd(&hist[l0], __popcll(same));
            pending &= ~same;
        }
        ph[r] = 0;
        bool emit = false;
        if (i + 1 < L)
        {
            const Key a = K[i], b = K[i + 1];
            const int cp = clz64(a ^ b) - 1;
            unsigned level = unsigned(cp) / bits;
            if (level > maxLevel) level = maxLevel;
            const unsigned slot = unsigned(b >> (63 - bits * (level + 1))) & ((1u << bits) - 1);
            if (slot == 1)
            {
                emit = true;
                ph[r] = placeholderKey(b & ~(viewRange(level, bits) - 1), int(bits * level));
            }
        }
        emitRows |= emit ? 1u << r : 0u;
        const std::uint64_t bal = __ballot(emit);
        if (lane == 0) waveCount[r][wave] = unsigned(__popcll(bal));
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned total = 0;
        for (int r = 0; r < kClassifyRows; ++r)
            for (int w = 0; w < kT / 64; ++w)
            {
                const unsigned c = waveCount[r][w];
                waveCount[r][w] = total;
                total += c;
            }
        blockBase = total ? atomicAdd(candCount, total) : 0u;
    }
    __syncthreads();
    const std::uint64_t below = (std::uint64_t(1) << lane) - 1;
#pragma unroll
    for (int r = 0; r < kClassifyRows; ++r)
    {
        const bool emit = (emitRows >> r) & 1u;
        const std::uint64_t bal = __ballot(emit);
        if (emit) cand[blockBase + waveCount[r][wave] + __popcll(bal & below)] = ph[r];
    }
    if (threadIdx.x < 64 && hist[threadIdx.x]) atomicAdd(&leavesPerLevel[threadIdx.x], hist[threadIdx.x]);
}

// buildLevelMetadataKernel (tree_view.cu:211-275), one wave: lane l holds
// level l. The node counts per level follow from the deepest level up
// (internal[l] = nodes[l+1] / children, nodes[l] = leaves[l] + internal[l]),
// one lane over LDS; the offsets are prefix sums across the wave.
__global__ __launch_bounds__(64) void metadataKernel(const int *leavesPerLevel, int *internalPer, int *levelOffset,
                                                     int *offInternal, std::uint8_t *maxDepth, int cap,
                                                     const DevState *st, unsigned bits)
{
    __shared__ int leaves[64], internal[64], nodes[64];
    if (activeLeaves(cap, st) == 0) return;
    const unsigned kMax = 63u / bits;
    const int children = 1 << bits;
    const unsigned l = threadIdx.x;
    const int lv = l <= kMax ? leavesPerLevel[l] : 0;
    leaves[l] = lv;
    internal[l] = 0;
    nodes[l] = 0;
    const std::uint64_t nonEmpty = __ballot(lv != 0);
    const unsigned maxLevel = nonEmpty ? 63u - unsigned(__clzll((long long)nonEmpty)) : 0u;
    __syncthreads();
    if (l == 0)
    {
        *maxDepth = std::uint8_t(maxLevel);
        nodes[maxLevel] = leaves[maxLevel];
        for (int k = int(maxLevel) - 1; k >= 0; --k)
        {
            internal[k] = nodes[k + 1] / children;
            nodes[k] = leaves[k] + internal[k];
        }
    }
    __syncthreads();
    int n = nodes[l], in = internal[l];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const int tn = __shfl_up(n, o, 64), ti = __shfl_up(in, o, 64);
        if (l >= unsigned(o)) n += tn, in += ti;
    }
    if (l <= kMax)
    {
        internalPer[l] = internal[l];
        levelOffset[l + 1] = n;
        offInternal[l + 1] = in;
    }
    if (l == 0) levelOffset[0] = offInternal[0] = 0;
}

__global__ __launch_bounds__(kT) void layoutKeysKernel(const Key *S, std::size_t icap, const DevState *st,
                                                       unsigned bits, Key *layoutKey, std::uint32_t *layoutVal)
{
    const std::size_t idx = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= icap) return;
    const int I = activeInternals(int(icap), st, bits);
    if (idx < std::size_t(I))
    {
        const Key ph = S[idx];
        const unsigned lvl = phLevel(ph, bits);
        layoutKey[idx] = (Key(1) << (bits * lvl)) | reversedPath(ph, lvl, bits);
    }
    else
        layoutKey[idx] = kInvalidKey;
    layoutVal[idx] = std::uint32_t(idx);
}

__global__ __launch_bounds__(kT) void materializeKernel(const Key *S, const std::uint32_t *layoutVal,
                                                        std::size_t icap, const DevState *st, unsigned bits,
                                                        const int *offInternal, int *mBySfc)
{
    const int li = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (li >= activeInternals(int(icap), st, bits)) return;
    const std::uint32_t sp = layoutVal[li];
    const Key ph = S[sp];
    const unsigned lvl = phLevel(ph, bits);
    mBySfc[sp] = li - offInternal[lvl];
}

__device__ inline int finalNodeIndex(Key ph, unsigned lvl, unsigned bits, const int *levelOffset,
                                     const int *offInternal, const int *internalPer, const Key *S, const int *mBySfc)
{
    if (lvl == 0) return 0;
    const Key parent = ph >> bits;
    const int b = offInternal[lvl - 1], e = offInternal[lvl];
    const int pos = b + int(lowerBound(S + b, std::size_t(e - b), parent));
    const unsigned slot = unsigned(ph) & ((1u << bits) - 1);
    return levelOffset[lvl] + int(slot) * internalPer[lvl - 1] + mBySfc[pos];
}

struct Rows
{
    Key *box;
    std::uint8_t *split, *depth;
    int *parent, *child;
    std::uint32_t *begin, *counts;
    int *leafToNode;
};

// An internal node's particles are those of its leaves [first, end) (its
// start and end keys are leaf boundaries): begin = leafBegin[first], count =
// leafBegin[end] - begin, with leafBegin[L] = the active count. Equal to
// lower_bound(P, start) and lower_bound(P, end) - lower_bound(P, start).
// Threads take the internal nodes in S order (by level, key order inside a
// level), so a wave's searches in S and K walk neighbouring keys; the rows
// land at the layout positions.
__global__ __launch_bounds__(kT) void internalRowsKernel(std::size_t icap, const DevState *st, unsigned bits,
                                                         const int *levelOffset, const int *offInternal,
                                                         const int *internalPer, const Key *S, const int *mBySfc,
                                                         const Key *K0, const Key *K1, const std::uint32_t *leafBegin,
                                                         int withCounts, Rows r)
{
    const int sp = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (sp >= activeInternals(int(icap), st, bits)) return;
    const Key ph = S[sp];
    const unsigned lvl = phLevel(ph, bits);
    const Key start = phStart(ph, lvl, bits);
    const int node = finalNodeIndex(ph, lvl, bits, levelOffset, offInternal, internalPer, S, mBySfc);
    r.box[node] = ph;
    r.depth[node] = std::uint8_t(lvl);
    r.split[node] = 1;
    if (lvl == 0) r.parent[node] = -1;
    const int m = mBySfc[sp];
    const int M = internalPer[lvl];
    r.child[node] = levelOffset[lvl + 1] + m;
    for (int s = 0; s < (1 << bits); ++s) r.parent[levelOffset[lvl + 1] + s * M + m] = node;
    if (withCounts)
    {
        const Key *K = st->activeBuffer ? K1 : K0;
        const std::size_t L1 = std::size_t(st->numLeaves) + 1;
        const std::size_t f = lowerBound(K, L1, start);
        const std::size_t e = lowerBound(K, L1, start + viewRange(lvl, bits));
        const std::uint32_t b = leafBegin[f];
        r.begin[node] = b;
        r.counts[node] = leafBegin[e] - b;
        if (lvl == 0) r.counts[st->numNodes] = leafBegin[e] - b;
    }
}

// Particles per block of the particle -> leaf-node mapping.
constexpr int kMapChunk = kT * 8;

// Exclusive scan of N over the live leaves (chained scan, 2048 per tile).
// With chunkLeaf, each leaf t with particles [b, e) also records itself as
// the leaf of every mapping chunk that starts inside it: chunkLeaf[c] = t
// for c * kMapChunk in [b, e).
__global__ __launch_bounds__(kT) void leafBeginKernel(const std::uint32_t *N0, const std::uint32_t *N1, int cap,
                                                      const DevState *st, std::uint32_t *out, std::uint64_t *status,
                                                      std::uint32_t *ticket, std::uint32_t *chunkLeaf)
{
    constexpr int kItems = 8;
    __shared__ std::uint32_t tileShared;
    __shared__ std::uint64_t wsum[kT / 64];
    __shared__ std::uint64_t prefixShared;
    if (threadIdx.x == 0) tileShared = atomicAdd(ticket, 1u);
    __syncthreads();
    const int L = activeLeaves(cap, st);
    const long base = long(tileShared) * kT * kItems;
    if (base > L || L == 0) return;
    const std::uint32_t *N = st->activeBuffer ? N1 : N0;
    const long first = base + long(threadIdx.x) * kItems;
    std::uint32_t v[kItems];
    std::uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        v[k] = first + k < L ? N[first + k] : 0u;
        sum += v[k];
    }
    const unsigned lane = __lane_id();
    const int wave = threadIdx.x >> 6;
    std::uint64_t inc = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const std::uint64_t t = __shfl_up(inc, o, 64);
        if (lane >= unsigned(o)) inc += t;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    std::uint64_t before = 0, tileSum = 0;
    for (int w = 0; w < kT / 64; ++w)
    {
        if (w < wave) before += wsum[w];
        tileSum += wsum[w];
    }
    if (threadIdx.x == 0)
    {
        const std::size_t tile = tileShared;
        std::uint64_t *me = status + tile;
        std::uint64_t acc = 0;
        if (tile == 0)
            stRel64(me, kPre | tileSum);
        else
        {
            stRel64(me, kAgg | tileSum);
            std::size_t j = tile - 1;
            for (;;)
            {
                const std::uint64_t w = ldAcq64(status + j);
                const std::uint64_t f = w & kFlagMask;
                if (!f)
                {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                acc += w & kSumMask;
                if (f == kPre) break;
                --j;
            }
            stRel64(me, kPre | (acc + tileSum));
        }
        prefixShared = acc;
    }
    __syncthreads();
    std::uint64_t run = prefixShared + before + inc - sum;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        if (first + k <= L) out[first + k] = std::uint32_t(run); // [L]: the total
        if (chunkLeaf && first + k < L)
            for (std::uint64_t c = (run + kMapChunk - 1) / kMapChunk; c * kMapChunk < run + v[k]; ++c)
                chunkLeaf[c] = std::uint32_t(first + k);
        run += v[k];
    }
}

__global__ void clearScanKernel(std::uint64_t *status, std::size_t words, std::uint32_t *ticket)
{
    const std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == 0) *ticket = 0;
    for (std::size_t j = i; j < words; j += std::size_t(gridDim.x) * blockDim.x) status[j] = 0;
}

__global__ __launch_bounds__(kT) void leafRowsKernel(const Key *K0, const Key *K1, const std::uint32_t *N0,
                                                     const std::uint32_t *N1, int cap, const DevState *st,
                                                     unsigned bits, const std::uint8_t *leafLevel,
                                                     const int *levelOffset, const int *offInternal,
                                                     const int *internalPer, const Key *S, const int *mBySfc,
                                                     const std::uint32_t *leafBegin, int withCounts, Rows r)
{
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= activeLeaves(cap, st)) return;
    const Key *K = st->activeBuffer ? K1 : K0;
    const std::uint32_t *N = st->activeBuffer ? N1 : N0;
    const unsigned lvl = leafLevel[i];
    const Key ph = placeholderKey(K[i], int(bits * lvl));
    const int node = finalNodeIndex(ph, lvl, bits, levelOffset, offInternal, internalPer, S, mBySfc);
    r.box[node] = ph;
    r.depth[node] = std::uint8_t(lvl);
    r.split[node] = 0;
    r.child[node] = -1;
    r.leafToNode[i] = node;
    if (lvl == 0) r.parent[node] = -1;
    if (withCounts)
    {
        r.begin[node] = leafBegin[i];
        r.counts[node] = N[i];
        if (lvl == 0) r.counts[st->numNodes] = N[i];
    }
}

// fillSfcParticleIndexKernel (tree_view.cu:542-562): particle p (sorted)
// maps to the view node of the leaf holding P[p], or to 2^64-1 when P[p] is
// escaped. The leaves partition the active particles [0, A) in order: leaf
// t holds [leafBegin[t], leafBegin[t+1]), so p's leaf is the last t with
// leafBegin[t] <= p. Each block takes kMapChunk consecutive particles; the
// leaves [t0, t1] they span come from chunkLeaf (the leaf holding the
// chunk's first particle, and the one holding the next chunk's, or the
// last leaf); it stages those begins in LDS when they fit (else searches
// them in place), and writes its particles in order; block work is fixed
// whatever the leaf sizes.
constexpr int kMapStage = 2048;
__global__ __launch_bounds__(kT) void mappingKernel(std::size_t n, int cap, const DevState *st,
                                                    const std::uint32_t *leafBegin, const std::uint32_t *chunkLeaf,
                                                    const int *leafToNode, Key *out)
{
    __shared__ std::uint32_t begins[kMapStage];
    __shared__ std::uint32_t span[2];
    const int L = activeLeaves(cap, st);
    if (L == 0) return;
    const std::size_t c0 = std::size_t(blockIdx.x) * kMapChunk;
    if (c0 >= n) return;
    const std::size_t cEnd = c0 + kMapChunk < n ? c0 + kMapChunk : n;
    const std::size_t active = leafBegin[L];
    const std::size_t c1 = cEnd < active ? cEnd : active;
    if (c0 < c1)
    {
        if (threadIdx.x == 0)
        {
            // c0 < active: a leaf holds c0; c1 < active: a leaf holds c1 = the next chunk's start
            span[0] = chunkLeaf[blockIdx.x];
            span[1] = c1 < active ? chunkLeaf[blockIdx.x + 1] : std::uint32_t(L - 1);
        }
        __syncthreads();
        const std::uint32_t t0 = span[0], m = span[1] - t0 + 1;
        const bool staged = m <= std::uint32_t(kMapStage);
        if (staged)
            for (std::uint32_t j = threadIdx.x; j < m; j += kT) begins[j] = leafBegin[t0 + j];
        __syncthreads();
        const std::uint32_t *B = staged ? begins : leafBegin + t0;
        for (std::size_t p = c0 + threadIdx.x; p < c1; p += kT)
        {
            // last t in [0, m) with B[t] <= p (B[0] <= c0 <= p)
            std::uint32_t lo = 0, len = m;
            while (len > 1)
            {
                const std::uint32_t half = len >> 1;
                if (B[lo + half] <= std::uint32_t(p)) lo += half;
                len -= half;
            }
            out[p] = Key(std::int64_t(leafToNode[t0 + lo]));
        }
    }
    for (std::size_t p = c1 + threadIdx.x; p < cEnd; p += kT) out[p] = kInvalidKey;
}
} // namespace

struct ViewEngine::Impl
{
    VecShape shape[kOuts];
    DevBuf<Key> box, mapping;
    DevBuf<std::uint8_t> split, depth;
    DevBuf<int> parent, child, levelOffset;
    DevBuf<std::uint32_t> begin, counts;
    // scratch
    DevBuf<int> leafToNode, leavesPerLevel, internalPer, offInternal, mBySfc;
    DevBuf<std::uint8_t> leafLevel;
    DevBuf<Key> cand, layoutKey;
    DevBuf<std::uint32_t> layoutVal, leafBegin, candCount, scanTicket;
    DevBuf<std::uint32_t> chunkLeaf; // per mapping chunk: the leaf holding its first particle
    DevBuf<std::uint64_t> scanStatus;
    SortWorkspace sort;
};

ViewEngine::ViewEngine(unsigned levelBits) : impl_(new Impl), bits_(levelBits), maxLevel_(63u / levelBits) {}
ViewEngine::~ViewEngine()
{
    delete impl_;
    poolRelease(depth_);
}

void ViewEngine::prepare(std::size_t leafCapacity, std::size_t numParticles, bool structure, bool counts,
                         bool mapping, hipStream_t s)
{
    if (!structure && !counts && !mapping) return;
    leafCap_ = leafCap_ > leafCapacity ? leafCap_ : leafCapacity;
    particleCap_ = particleCap_ > numParticles ? particleCap_ : numParticles;
    prepStructure_ |= structure;
    prepCounts_ |= counts;
    prepMapping_ |= mapping;
    const std::size_t internals = (leafCap_ - 1) / ((1u << bits_) - 1);
    const std::size_t nodes = leafCap_ + internals;
    Impl &m = *impl_;
    scratch_ = true;
    if (!depth_)
    {
        depth_ = static_cast<std::uint8_t *>(poolAlloc(1, s));
        AT_HIP(hipMemsetAsync(depth_, 0, 1, s));
    }
    auto grow = [&](Out o, std::size_t n) { m.shape[o].resizeHeadroom(n); };
    grow(BoxIndex, nodes);
    grow(HasSplit, nodes);
    grow(Depth, nodes);
    grow(Parent, nodes);
    grow(Child, nodes);
    grow(LevelOffset, maxLevel_ + 2);
    if (prepCounts_)
    {
        grow(Begin, nodes);
        grow(Counts, nodes + 1);
    }
    if (prepMapping_) grow(Mapping, particleCap_);
    m.box.ensure(m.shape[BoxIndex].cap, m.box.cap, s);
    m.split.ensure(m.shape[HasSplit].cap, m.split.cap, s);
    m.depth.ensure(m.shape[Depth].cap, m.depth.cap, s);
    m.parent.ensure(m.shape[Parent].cap, m.parent.cap, s);
    m.child.ensure(m.shape[Child].cap, m.child.cap, s);
    m.levelOffset.ensure(m.shape[LevelOffset].cap, m.levelOffset.cap, s);
    if (prepCounts_)
    {
        m.begin.ensure(m.shape[Begin].cap, m.begin.cap, s);
        m.counts.ensure(m.shape[Counts].cap, m.counts.cap, s);
    }
    if (prepMapping_)
    {
        m.mapping.ensure(m.shape[Mapping].cap, m.mapping.cap, s);
        m.chunkLeaf.ensure(particleCap_ / kMapChunk + 2, 0, s);
    }
    m.leafToNode.ensure(leafCap_, m.leafToNode.cap, s);
    m.leavesPerLevel.ensure(64, 0, s);
    m.internalPer.ensure(64, 0, s);
    m.offInternal.ensure(66, 0, s);
    m.leafLevel.ensure(leafCap_, 0, s);
    m.leafBegin.ensure(leafCap_ + 1, 0, s);
    const std::size_t ic = internals ? internals : 1;
    m.cand.ensure(ic, 0, s);
    m.layoutKey.ensure(ic, 0, s);
    m.layoutVal.ensure(ic, 0, s);
    m.mBySfc.ensure(ic, 0, s);
    m.candCount.ensure(1, 0, s);
    m.scanTicket.ensure(1, 0, s);
    m.scanStatus.ensure(leafCap_ / (kT * 8) + 2, 0, s);
    if (internals > 1) m.sort.allocate(internals, true, s);
}

void ViewEngine::syncSizes(std::size_t numLeaves, std::size_t numActive)
{
    if (!scratch_ || numLeaves > leafCap_) return;
    numNodes_ = numLeaves + (numLeaves - 1) / ((1u << bits_) - 1);
    Impl &m = *impl_;
    for (Out o : {BoxIndex, HasSplit, Depth, Parent, Child}) m.shape[o].resize(numNodes_);
    if (prepCounts_)
    {
        m.shape[Begin].resize(numNodes_);
        m.shape[Counts].resize(numNodes_ + 1);
    }
    if (prepMapping_) m.shape[Mapping].resize(numActive);
}

void ViewEngine::build(hipStream_t s, Key *const K[2], std::uint32_t *const N[2], const Key *P, std::size_t n,
                       bool structure, bool counts, bool mapping, DevState *st)
{
    if (!structure && !counts && !mapping) return;
    if (!scratch_ || (counts && !prepCounts_) || (mapping && (!prepMapping_ || n > particleCap_)))
    {
        hipLaunchKernelGGL(flagResizeKernel, dim3(1), dim3(1), 0, s, st);
        return;
    }
    Impl &m = *impl_;
    prof::Scope scope("views", s);
    const int cap = int(leafCap_);
    const std::size_t icap = (leafCap_ - 1) / ((1u << bits_) - 1);
    const unsigned leafBlocks = ceilDiv(leafCap_, kT);
    Rows rows{m.box.p, m.split.p, m.depth.p, m.parent.p, m.child.p, counts ? m.begin.p : nullptr,
              counts ? m.counts.p : nullptr, m.leafToNode.p};
    const unsigned prepBlocks = ceilDiv(icap > 64 ? icap : 64, kT);
    hipLaunchKernelGGL(prepKernel, dim3(prepBlocks > 1024 ? 1024 : prepBlocks), dim3(kT), 0, s, m.leavesPerLevel.p,
                       int(maxLevel_ + 1), m.candCount.p, m.cand.p, icap);
    hipLaunchKernelGGL(classifyKernel, dim3(ceilDiv(leafCap_, kT * kClassifyRows)), dim3(kT), 0, s, K[0], K[1], st,
                       cap, bits_, m.leafLevel.p,
                       m.leavesPerLevel.p, m.cand.p, m.candCount.p);
    hipLaunchKernelGGL(metadataKernel, dim3(1), dim3(64), 0, s, m.leavesPerLevel.p, m.internalPer.p, m.levelOffset.p,
                       m.offInternal.p, depth_, cap, st, bits_);
    if (counts || mapping)
    {
        const std::size_t words = leafCap_ / (kT * 8) + 2;
        hipLaunchKernelGGL(clearScanKernel, dim3(ceilDiv(words, kT)), dim3(kT), 0, s, m.scanStatus.p, words,
                           m.scanTicket.p);
        hipLaunchKernelGGL(leafBeginKernel, dim3(unsigned(leafCap_ / (kT * 8) + 1)), dim3(kT), 0, s, N[0], N[1], cap,
                           st, m.leafBegin.p, m.scanStatus.p, m.scanTicket.p, mapping ? m.chunkLeaf.p : nullptr);
    }
    if (icap > 0)
    {
        if (icap > 1) sortKeys(s, m.cand.p, icap, m.sort);
        const unsigned ib = ceilDiv(icap, kT);
        hipLaunchKernelGGL(layoutKeysKernel, dim3(ib), dim3(kT), 0, s, m.cand.p, icap, st, bits_, m.layoutKey.p,
                           m.layoutVal.p);
        if (icap > 1) sortPairs(s, m.layoutKey.p, m.layoutVal.p, icap, m.sort, false, false);
        hipLaunchKernelGGL(materializeKernel, dim3(ib), dim3(kT), 0, s, m.cand.p, m.layoutVal.p, icap, st, bits_,
                           m.offInternal.p, m.mBySfc.p);
        hipLaunchKernelGGL(internalRowsKernel, dim3(ib), dim3(kT), 0, s, icap, st, bits_, m.levelOffset.p, m.offInternal.p, m.internalPer.p, m.cand.p, m.mBySfc.p, K[0], K[1],
                           m.leafBegin.p, counts ? 1 : 0, rows);
    }
    hipLaunchKernelGGL(leafRowsKernel, dim3(leafBlocks), dim3(kT), 0, s, K[0], K[1], N[0], N[1], cap, st, bits_,
                       m.leafLevel.p, m.levelOffset.p, m.offInternal.p, m.internalPer.p, m.cand.p, m.mBySfc.p,
                       m.leafBegin.p, counts ? 1 : 0, rows);
    if (mapping && n)
        hipLaunchKernelGGL(mappingKernel, dim3(ceilDiv(n, kMapChunk)), dim3(kT), 0, s, n, cap, st, m.leafBegin.p,
                           m.chunkLeaf.p, m.leafToNode.p, m.mapping.p);
    AT_HIP(hipGetLastError());
}

unsigned ViewEngine::maxAchievedDepth(hipStream_t s) const
{
    if (!depth_) return 0;
    std::uint8_t v = 0;
    syncStream(s);
    AT_HIP(hipMemcpy(&v, depth_, 1, hipMemcpyDeviceToHost));
    return v;
}

void ViewEngine::buffer(Out o, const void **d, std::size_t *size, std::size_t *cap) const
{
    const Impl &m = *impl_;
    *size = m.shape[o].size;
    *cap = m.shape[o].cap;
    switch (o)
    {
    case BoxIndex: *d = m.box.p; break;
    case HasSplit: *d = m.split.p; break;
    case Depth: *d = m.depth.p; break;
    case Parent: *d = m.parent.p; break;
    case Child: *d = m.child.p; break;
    case LevelOffset: *d = m.levelOffset.p; break;
    case Begin: *d = m.begin.p; break;
    case Counts: *d = m.counts.p; break;
    case Mapping: *d = m.mapping.p; break;
    default: *d = nullptr;
    }
}
} // namespace at
