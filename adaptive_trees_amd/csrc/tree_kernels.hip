// tree_kernels.hip -- octree and binary-tree rebalance kernels for gfx950.
//
// One split/merge pass is three launches that read every size from the
// device state, so a pass is graph-capturable and grids are sized by
// capacity:
//   decideScan  per-leaf decision (octree_rebalance.cuh:174-249 /
//               kdtree3d_rebalance.cuh:61-81), block scan and a chained scan
//               across tiles in the same kernel; the tile that holds the
//               sentinel index applies the capacity rule
//               (octree_rebalance.cuh:1211-1232).
//   process     scatter into the inactive A/B half (octree_rebalance.cuh:
//               253-365, 1234-1265 / kdtree3d_rebalance.cuh:103-117), then a
//               one-block commit that flips the halves and clears the scan
//               status for the next pass.
//   refresh     leaf counts, or near-field counts of dirty leaves and
//               dirty sibling groups from a compacted work list.
#include "tree_kernels.hpp"

#include "profile.hpp"

#include <algorithm>

namespace at
{
namespace
{
constexpr std::uint64_t kAgg = std::uint64_t(1) << 62;
constexpr std::uint64_t kPre = std::uint64_t(2) << 62;
constexpr std::uint64_t kChg = std::uint64_t(1) << 61;
constexpr std::uint64_t kSumMask = kChg - 1;
constexpr std::uint64_t kFlagMask = std::uint64_t(3) << 62;
constexpr int kThreads = 256;
constexpr int kItems = kPassTile / kThreads;
// single-tree decision scans: 1024 x kItems leaves per tile, so a pass draws
// a quarter of the tile tickets and look-back steps of kPassTile tiles
constexpr int kDecideThreads = 1024;
inline std::size_t decideTiles(std::size_t cap) { return cap / (std::size_t(kDecideThreads) * kItems) + 1; }

// Chained scan across tiles (one status word per tile: flag, change bit, sum).
// A tile publishes its aggregate, looks back for its exclusive prefix and
// publishes its inclusive prefix. The look-back is done by one whole wave:
// each step loads the status words of the kLookWindows * 64 nearest tiles not
// yet accounted for (window w, lane k: tile j - 64 w - k), all loads in
// flight at once, then consumes the windows nearest first. It stops at the
// nearest tile with a prefix once every nearer one has published its
// aggregate; a window with a tile not yet published ends the step (the
// windows before it are kept). Positions before tile 0 read as an empty
// prefix. Called by all 64 lanes of one wave.
constexpr int kLookWindows = 4;
__device__ void chainedScanWave(std::uint64_t *status, std::size_t tile, std::uint64_t tileSum, bool tileChg,
                                std::uint64_t &exclusive, bool &exclChg)
{
    const unsigned lane = __lane_id();
    std::uint64_t *me = status + tile;
    if (tile == 0)
    {
        if (lane == 0) stRel64(me, kPre | (tileChg ? kChg : 0) | tileSum);
        exclusive = 0;
        exclChg = false;
        return;
    }
    if (lane == 0) stRel64(me, kAgg | (tileChg ? kChg : 0) | tileSum);
    std::uint64_t acc = 0;
    bool chg = false;
    std::size_t j = tile - 1;
    for (;;)
    {
        std::uint64_t v[kLookWindows];
#pragma unroll
        for (int w = 0; w < kLookWindows; ++w)
        {
            const std::size_t back = std::size_t(w) * 64 + lane;
            v[w] = j >= back ? ldAcq64(status + (j - back)) : kPre;
        }
        int used = 0;
        bool done = false;
#pragma unroll
        for (int w = 0; w < kLookWindows; ++w)
        {
            const std::uint64_t f = v[w] & kFlagMask;
            const std::uint64_t notReady = __ballot(f == 0);
            const std::uint64_t pre = __ballot(f == kPre);
            const int stop = pre ? __ffsll((long long)pre) - 1 : 63; // nearest prefix, or the whole window
            const std::uint64_t need = stop == 63 ? ~std::uint64_t(0) : (std::uint64_t(1) << (stop + 1)) - 1;
            if (notReady & need) break; // wave-uniform
            unsigned long long part = int(lane) <= stop ? (v[w] & kSumMask) : 0ull;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) part += __shfl_xor(part, o, 64);
            acc += part;
            chg |= __ballot(int(lane) <= stop && (v[w] & kChg) != 0) != 0;
            ++used;
            if (pre)
            {
                done = true;
                break;
            }
        }
        if (done) break;
        // a consumed window without a prefix lies wholly above tile 0 (tile 0
        // always holds one), so j stays >= 0
        j -= std::size_t(used) * 64;
        if (used < kLookWindows) __builtin_amdgcn_s_sleep(1);
    }
    if (lane == 0) stRel64(me, kPre | ((chg || tileChg) ? kChg : 0) | (acc + tileSum));
    exclusive = acc;
    exclChg = chg;
}

// Block-wide exclusive scan of a 64-bit value and OR of a flag.
template <int THREADS>
__device__ std::uint64_t blockScan64(std::uint64_t v, bool flag, std::uint64_t *wsum, int *wflag,
                                     std::uint64_t &total, bool &anyFlag)
{
    const unsigned lane = __lane_id();
    const int wave = threadIdx.x >> 6;
    std::uint64_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const std::uint64_t t = __shfl_up(inc, o, 64);
        if (lane >= unsigned(o)) inc += t;
    }
    const std::uint64_t fb = __ballot(flag);
    if (lane == 63)
    {
        wsum[wave] = inc;
        wflag[wave] = fb != 0;
    }
    __syncthreads();
    std::uint64_t before = 0;
    total = 0;
    anyFlag = false;
#pragma unroll
    for (int w = 0; w < THREADS / 64; ++w)
    {
        if (w < wave) before += wsum[w];
        total += wsum[w];
        anyFlag |= wflag[w] != 0;
    }
    return before + inc - v;
}

// ---- octree decision (calculateNodeOp, octree_rebalance.cuh:174-249) --------
__device__ inline int siblingIndex(const Key *tree, int numNodes, int i, unsigned &level)
{
    const Key node = tree[i];
    level = treeLevel(tree[i + 1] - node);
    if (level == 0) return -1;
    const int s = int(octalDigit(node, level));
    const int first = i - s;
    if (first < 0 || first > numNodes - 8) return -1;
    return tree[first + 8] == tree[first] + nodeRange(level - 1) ? s : -1;
}

struct OctDecider
{
    const Key *K[2];
    const std::uint32_t *N[2];
    const std::uint32_t *NF[2];
    const std::uint8_t *gcm[2];
    unsigned bucket;
    bool nf;

    // the active half by selects (an index into the pairs would put them in
    // per-lane scratch memory)
    __device__ int operator()(int ab, int i, int L) const
    {
        const Key *tree = ab ? K[1] : K[0];
        const std::uint32_t *n = ab ? N[1] : N[0];
        unsigned level = 0;
        const int s = siblingIndex(tree, L, i, level);
        if (!nf)
        {
            if (s > 0)
            {
                const std::uint32_t *g = n + i - s;
                std::uint64_t parent = 0;
#pragma unroll
                for (int k = 0; k < 8; ++k) parent += g[k];
                if (parent <= std::uint64_t(bucket)) return 0;
            }
            const std::uint32_t c = n[i];
            // 32-bit threshold products, as the reference computes them
            if (c > bucket * 512u && level + 3 < kOctLevels) return 4096;
            if (c > bucket * 64u && level + 2 < kOctLevels) return 512;
            if (c > bucket * 8u && level + 1 < kOctLevels) return 64;
            if (c > bucket && level < kOctLevels) return 8;
            return 1;
        }
        if (s >= 0)
        {
            const int first = i - s;
            if (first + 8 <= L && (ab ? gcm[1] : gcm[0])[first]) return s == 0 ? 1 : 0;
        }
        const std::uint64_t c = (ab ? NF[1] : NF[0])[i];
        const std::uint64_t b = bucket;
        if (c > b * 262144u && level + 3 < kOctLevels) return 4096;
        if (c > b * 4096u && level + 2 < kOctLevels) return 512;
        if (c > b * 64u && level + 1 < kOctLevels) return 64;
        if (c > b && level < kOctLevels) return 8;
        return 1;
    }
};

// ---- binary decision (kdtree3d_rebalance.cuh:61-81, binary_node.cuh:91-98) ---
__device__ inline bool binSiblings(const Key *tree, int count, int first)
{
    if (first < 0 || first + 1 >= count) return false;
    const Key r = tree[first + 1] - tree[first];
    return !(tree[first] & r) && tree[first + 2] - tree[first + 1] == r;
}

struct BinDecider
{
    const Key *K[2];
    const std::uint32_t *N[2];
    unsigned bucket;
    unsigned maxDepth;

    __device__ int operator()(int ab, int i, int L) const
    {
        const Key *tree = ab ? K[1] : K[0];
        const std::uint32_t *c = ab ? N[1] : N[0];
        const bool merge = i > 0 && binSiblings(tree, L, i - 1) &&
                           std::uint64_t(c[i - 1]) + c[i] <= std::uint64_t(bucket);
        if (merge) return 0;
        return (c[i] > bucket && unsigned(clz64(tree[i + 1] - tree[i])) < maxDepth) ? 2 : 1;
    }
};

// THREADS x kItems leaves per tile (kDecideThreads for single trees, kThreads
// in the batch kernels)
template <int THREADS, class Dec>
__device__ void decideScanBody(const Dec &dec, DevState *st, long capacity, std::int64_t *offs,
                               std::uint64_t *status, std::uint32_t *ticket)
{
    constexpr long kTile = long(THREADS) * kItems;
    __shared__ std::uint32_t tileShared;
    __shared__ std::uint64_t wsum[THREADS / 64];
    __shared__ int wflag[THREADS / 64];
    __shared__ std::uint64_t prefixShared;
    __shared__ int prefixChgShared;
    const int tid = threadIdx.x;
    // A pass after convergence decides nothing and its offsets are never
    // read (process and refresh return on the flag; the commit clears the
    // status words and ticket). Only update passes reach this: a build stops
    // at the pass that converges. A block that reads the flag set by the last
    // tile of this same launch would have drawn a ticket past that tile.
    if (st->converged) return;
    if (tid == 0) tileShared = atomicAdd(ticket, 1u);
    __syncthreads();
    const long tile = tileShared;
    const long L = st->numLeaves;
    const long base = tile * kTile;
    if (base > L) return; // tiles past the sentinel index hold no operations
    const bool skip = st->converged || st->needsResize;
    const int ab = st->activeBuffer;
    // decisions in striped order (leaf base + k * THREADS + tid: neighbouring
    // lanes read neighbouring leaves), staged in LDS; the scan takes each
    // thread's run of kItems consecutive leaves from there
    __shared__ std::uint16_t opStage[THREADS * kItems];
    __shared__ std::uint32_t prefStage[THREADS * kItems];
    bool chg = false;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        const long i = base + long(k) * THREADS + tid;
        int o = 0;
        if (!skip && i < L)
        {
            o = dec(ab, int(i), int(L));
            chg |= o != 1;
        }
        opStage[k * THREADS + tid] = std::uint16_t(o);
    }
    __syncthreads();
    std::uint32_t op[kItems];
    std::uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        op[k] = opStage[tid * kItems + k];
        sum += op[k];
    }
    std::uint64_t tileSum;
    bool tileChg;
    const std::uint64_t local = blockScan64<THREADS>(sum, chg, wsum, wflag, tileSum, tileChg);
    if (tid < 64)
    {
        std::uint64_t ex;
        bool exChg;
        chainedScanWave(status, std::size_t(tile), tileSum, tileChg, ex, exChg);
        if (tid == 0)
        {
            prefixShared = ex;
            prefixChgShared = (exChg || tileChg) ? 1 : 0;
        }
    }
    // tile-local exclusive offsets of the thread's run, read back in striped order
    {
        std::uint32_t r = std::uint32_t(local);
#pragma unroll
        for (int k = 0; k < kItems; ++k)
        {
            prefStage[tid * kItems + k] = r;
            r += op[k];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        const long i = base + long(k) * THREADS + tid;
        const std::uint64_t run = prefixShared + prefStage[k * THREADS + tid];
        if (i <= L) offs[i] = std::int64_t(run);
        if (i == L)
        {
            // capacity rule of the pass (graphCheckCapacityKernel)
            const bool changedAll = prefixChgShared != 0;
            if (!st->converged)
            {
                if (st->needsResize || !changedAll) st->converged = 1;
                else if (long(run) > capacity)
                {
                    st->needsResize = 1;
                    st->converged = 1;
                }
                else
                    st->newNumLeaves = int(run);
            }
            st->changed = changedAll ? 1 : 0;
        }
    }
}

template <class Dec>
__global__ __launch_bounds__(kDecideThreads) void decideScanKernel(Dec dec, DevState *st, long capacity,
                                                             std::int64_t *offs, std::uint64_t *status,
                                                             std::uint32_t *ticket)
{
    decideScanBody<kDecideThreads>(dec, st, capacity, offs, status, ticket);
}

// ---- octree scatter (processNode) -------------------------------------------
// One thread per old leaf writes a kept leaf and the children of splits into
// 8; the children of the level-jumping splits (64, 512 or 4096) are written
// by the whole wave, one such leaf after the other, so a pass that splits a
// few leaves into thousands of children each does not run on a few lanes.
__device__ void octProcessBody(const TreeDev &t, bool nf, unsigned block)
{
    DevState *st = t.st;
    if (st->converged) return;
    const int L = st->numLeaves;
    const int i = int(block * blockDim.x + threadIdx.x);
    const int ab = st->activeBuffer;
    const Key *tree = t.K[ab];
    Key *out = t.K[ab ^ 1];
    const std::int64_t *ops = t.offs;
    if (i == 0) out[st->newNumLeaves] = nodeRange(0);
    const int code = i < L ? int(ops[i + 1] - ops[i]) : 0;
    const Key node = i < L && code != 0 ? tree[i] : Key(0);
    const std::int64_t dst = i < L ? ops[i] : 0;
    if (code == 1)
    {
        out[dst] = node;
        if (!nf && t.countDirty)
        {
            // kept leaf: same count; a merge leader absorbs its 7 siblings (op 0)
            std::uint32_t c = t.N[ab][i];
            if (i + 1 < L && ops[i + 2] == ops[i + 1])
                for (int k = 1; k < 8; ++k) c += t.N[ab][i + k];
            t.N[ab ^ 1][dst] = c;
            t.countDirty[dst] = 0;
        }
        if (nf)
        {
            const bool leader = i + 1 < L && ops[i + 2] - ops[i + 1] == 0;
            if (leader)
            {
                t.nfDirty[dst] = 1;
                t.groupDirty[dst] = 1;
            }
            else
            {
                t.N[ab ^ 1][dst] = t.N[ab][i];
                t.NF[ab ^ 1][dst] = t.NF[ab][i];
                t.nfDirty[dst] = 0;
                unsigned sl = 0;
                const int s = siblingIndex(tree, L, i, sl);
                bool reusable = false;
                if (s == 0 && i + 8 <= L)
                {
                    reusable = true;
                    for (int k = 0; k < 8; ++k)
                        if (ops[i + k + 1] - ops[i + k] != 1) { reusable = false; break; }
                }
                if (reusable)
                {
                    t.gcm[ab ^ 1][dst] = t.gcm[ab][i];
                    t.groupDirty[dst] = 0;
                }
                else
                    t.groupDirty[dst] = 1;
            }
        }
    }
    Key r = 0;
    if (code > 1)
    {
        const unsigned level = treeLevel(tree[i + 1] - node);
        const unsigned diff = code == 8 ? 1u : (code == 64 ? 2u : (code == 512 ? 3u : 4u));
        r = nodeRange(level + diff);
    }
    if (code == 8)
        for (int s = 0; s < 8; ++s)
        {
            out[dst + s] = node + Key(s) * r;
            if (!nf && t.countDirty) t.countDirty[dst + s] = 1; // counted by the refresh
            if (nf)
            {
                t.nfDirty[dst + s] = 1;
                t.groupDirty[dst + s] = 1;
            }
        }
    // level jumps: every lane of the wave takes part (the lanes past L too)
    std::uint64_t big = __ballot(code > 8);
    const unsigned lane = __lane_id();
    while (big)
    {
        const int src = __ffsll(static_cast<unsigned long long>(big)) - 1;
        big &= big - 1;
        const int n = __shfl(code, src, 64);
        const Key first = __shfl(node, src, 64), step = __shfl(r, src, 64);
        const std::int64_t base = __shfl(dst, src, 64);
        for (int s = int(lane); s < n; s += 64)
        {
            out[base + s] = first + Key(s) * step;
            if (!nf && t.countDirty) t.countDirty[base + s] = 1;
            if (nf)
            {
                t.nfDirty[base + s] = 1;
                t.groupDirty[base + s] = 1;
            }
        }
    }
}

__global__ __launch_bounds__(kThreads) void octProcessKernel(TreeDev t, bool nf) { octProcessBody(t, nf, blockIdx.x); }

// ---- binary scatter (binaryProcessNodesKernel) --------------------------------
__device__ void binProcessBody(const TreeDev &t, unsigned block)
{
    DevState *st = t.st;
    if (st->converged) return;
    const int L = st->numLeaves;
    const int i = int(block * blockDim.x + threadIdx.x);
    if (i >= L) return;
    const int ab = st->activeBuffer;
    const Key *tree = t.K[ab];
    Key *out = t.K[ab ^ 1];
    if (i == 0) out[st->newNumLeaves] = kMaxKey;
    const int op = int(t.offs[i + 1] - t.offs[i]);
    if (!op) return;
    const std::int64_t dst = t.offs[i];
    // The new leaf counts follow from the old ones (exact counts of the same
    // P): a kept leaf keeps its count, a leaf whose right sibling merges into
    // it (op 0 on i+1) adds it, a split leaf searches its midpoint. Equal to
    // recounting every leaf (calculateNodeCount) after the pass.
    const std::uint32_t *nOld = t.N[ab];
    std::uint32_t *nNew = t.N[ab ^ 1];
    out[dst] = tree[i];
    if (op == 1)
    {
        const bool absorbs = i + 1 < L && t.offs[i + 2] == t.offs[i + 1];
        nNew[dst] = nOld[i] + (absorbs ? nOld[i + 1] : 0u);
        return;
    }
    const Key mid = tree[i] + (tree[i + 1] - tree[i]) / 2;
    out[dst + 1] = mid;
    std::size_t a, b;
    lowerBoundIndexedPair(t.P, t.n, t.Pidx, tree[i], mid, a, b);
    const std::uint32_t left = std::uint32_t(b - a);
    nNew[dst] = left;
    nNew[dst + 1] = nOld[i] - left;
}

__global__ __launch_bounds__(kThreads) void binProcessKernel(TreeDev t) { binProcessBody(t, blockIdx.x); }

// Commit the pass and clear the chained-scan status for the next one.
// passLog (device-driven batch builds): the new leaf count of each committed
// pass, for the host's mirror of the build's buffer capacities.
__device__ void commitBody(DevState *st, int binary, std::uint64_t *status, std::size_t words, std::uint32_t *ticket,
                           std::uint32_t *nfCount, unsigned block, unsigned blocks, std::uint32_t *passLog = nullptr)
{
    if (block == 0 && threadIdx.x == 0)
    {
        if (!st->converged)
        {
            if (passLog) passLog[st->passes] = std::uint32_t(st->newNumLeaves);
            st->numLeaves = st->newNumLeaves;
            if (binary) st->numNodes = (st->numLeaves - 1) * 2 + 1;
            st->activeBuffer ^= 1;
            st->passes += 1;
        }
        *ticket = 0;
        if (nfCount) nfCount[0] = nfCount[1] = 0;
    }
    for (std::size_t w = std::size_t(block) * blockDim.x + threadIdx.x; w < words; w += std::size_t(blocks) * blockDim.x)
        status[w] = 0;
}
__global__ void commitKernel(DevState *st, int binary, std::uint64_t *status, std::size_t words,
                             std::uint32_t *ticket, std::uint32_t *nfCount)
{
    commitBody(st, binary, status, words, ticket, nfCount, blockIdx.x, gridDim.x);
}

// Start of update(): resetRebalanceStateKernel (octree_rebalance.cuh:62-73).
__device__ void resetBody(DevState *st, const Key *P, std::size_t n, std::uint64_t *status, std::size_t words,
                          std::uint32_t *ticket, std::uint32_t *nfCount, unsigned block, unsigned blocks)
{
    if (block == 0 && threadIdx.x == 0)
    {
        st->changed = 0;
        st->converged = 0;
        st->passes = 0;
        st->numActiveParticles = n ? unsigned(lowerBound(P, n, kMaxKey)) : 0u;
        *ticket = 0;
        if (nfCount) nfCount[0] = nfCount[1] = 0;
    }
    for (std::size_t w = std::size_t(block) * blockDim.x + threadIdx.x; w < words; w += std::size_t(blocks) * blockDim.x)
        status[w] = 0;
}
__global__ void resetKernel(DevState *st, const Key *P, std::size_t n, std::uint64_t *status, std::size_t words,
                            std::uint32_t *ticket, std::uint32_t *nfCount)
{
    resetBody(st, P, n, status, words, ticket, nfCount, blockIdx.x, gridDim.x);
}

// initRootOctreeKernel + resetRebalanceState of the build (octree.cu:24-38, 159-163)
__device__ void octInitBody(const TreeDev &t, std::size_t words, unsigned block, unsigned blocks)
{
    if (block == 0 && threadIdx.x == 0)
    {
        Key *K = t.K[0];
        K[0] = 0;
        K[1] = nodeRange(0);
        const std::uint64_t active = t.n ? std::uint64_t(lowerBound(t.P, t.n, kMaxKey)) : 0;
        t.N[0][0] = unsigned(active);
        const std::uint64_t sq = active * active;
        if (t.NF[0]) t.NF[0][0] = sq > 0xffffffffull / 27ull ? 0xffffffffu : unsigned(27ull * sq);
        DevState s{};
        s.numLeaves = 1;
        s.newNumLeaves = 1;
        s.numActiveParticles = unsigned(active);
        *t.st = s;
        *t.scanTicket = 0;
        if (t.nfCount) t.nfCount[0] = t.nfCount[1] = 0;
    }
    for (std::size_t w = std::size_t(block) * blockDim.x + threadIdx.x; w < words; w += std::size_t(blocks) * blockDim.x)
        t.scanStatus[w] = 0;
}
__global__ void octInitKernel(TreeDev t, std::size_t words) { octInitBody(t, words, blockIdx.x, gridDim.x); }

// initBinaryRootKernel + refreshCounts (kdtree3d.cu:24-30, 139-146)
__device__ void binInitBody(const TreeDev &t, std::size_t words, unsigned block, unsigned blocks)
{
    if (block == 0 && threadIdx.x == 0)
    {
        t.K[0][0] = 0;
        t.K[0][1] = kMaxKey;
        const unsigned active = t.n ? unsigned(lowerBound(t.P, t.n, kMaxKey)) : 0u;
        DevState s{};
        s.numLeaves = 1;
        s.newNumLeaves = 1;
        s.numNodes = 1;
        s.numActiveParticles = active;
        *t.st = s;
        t.N[0][0] = active;
        *t.scanTicket = 0;
    }
    for (std::size_t w = std::size_t(block) * blockDim.x + threadIdx.x; w < words; w += std::size_t(blocks) * blockDim.x)
        t.scanStatus[w] = 0;
}
__global__ void binInitKernel(TreeDev t, std::size_t words) { binInitBody(t, words, blockIdx.x, gridDim.x); }

// N[i] = lb(P, K[i+1]) - lb(P, K[i]) over all slots (calculateNodeCount).
// Without `full`, leaves whose count the process step derived are skipped.
__device__ void leafCountsBody(const TreeDev &t, unsigned block, int full = 1)
{
    DevState *st = t.st;
    if (st->converged) return;
    const int L = st->numLeaves;
    const int i = int(block * blockDim.x + threadIdx.x);
    if (i >= L) return;
    if (!full && t.countDirty && !t.countDirty[i]) return;
    const int ab = st->activeBuffer;
    const Key *K = t.K[ab];
    std::size_t a, b;
    lowerBoundIndexedPair(t.P, t.n, t.Pidx, K[i], K[i + 1], a, b);
    const std::size_t c = b - a;
    t.N[ab][i] = c > 0xffffffffu ? 0xffffffffu : unsigned(c);
}
__global__ __launch_bounds__(kThreads) void leafCountsKernel(TreeDev t, int full)
{
    leafCountsBody(t, blockIdx.x, full);
}

// ---- near-field work list ------------------------------------------------------
__device__ inline bool dirtyLeader(const Key *tree, const std::uint8_t *gDirty, int L, int i)
{
    if (i + 8 > L) return false;
    if (gDirty && !gDirty[i]) return false;
    unsigned sl = 0;
    return siblingIndex(tree, L, i, sl) == 0 && sl > 0;
}

// Work lists of one refresh: a complete sibling group whose 8 leaves all need
// their own NF becomes one group item (its 8 stencils share a 4x4x4 block of
// cells); other dirty leaves are single items, dirty group leaders parent
// items. Items are appended with one atomic per list per block.
// Block `block` classifies ROWS rows of kThreads leaves, then takes its
// places in both lists with one atomic each (every block's count on the same
// two counters: with one row per block, 4.7M leaves made ~18K serialised
// atomics per counter and pass).
template <int ROWS>
__device__ void nfCollectBody(const TreeDev &t, int full, unsigned block)
{
    __shared__ std::uint32_t waveCnt[ROWS][kThreads / 64][2];
    __shared__ std::uint32_t blockBase[2];
    DevState *st = t.st;
    if (st->converged) return;
    const int L = st->numLeaves;
    const int ab = st->activeBuffer;
    const Key *tree = t.K[ab];
    const unsigned lane = __lane_id();
    const int wave = threadIdx.x >> 6;
    std::uint32_t flags = 0; // bits 4r..4r+3: own, lead, grp, grp with the parent's test of row r
#pragma unroll 1
    for (int r = 0; r < ROWS; ++r)
    {
        const int i = int((block * unsigned(ROWS) + unsigned(r)) * blockDim.x + threadIdx.x);
        bool own = false, lead = false, grp = false;
        if (i < L)
        {
            own = full || t.nfDirty[i];
            lead = dirtyLeader(tree, full ? nullptr : t.groupDirty, L, i);
            unsigned level = 0;
            const int s = own ? siblingIndex(tree, L, i, level) : -1;
            if (s >= 0)
            {
                const int g = i - s;
                bool all = true;
                if (!full)
                    for (int k = 0; k < 8; ++k) all = all && t.nfDirty[g + k] != 0;
                if (all)
                {
                    own = false;
                    grp = s == 0;
                }
            }
        }
        // a group item whose leader is also a dirty leader does the parent's
        // merge test too (its 27 parent-level cells are the ones the group's
        // own stencils are bracketed by)
        const bool grpLead = grp && lead;
        lead = lead && !grp;
        flags |= (std::uint32_t(own) | std::uint32_t(lead) << 1 | std::uint32_t(grp) << 2 |
                  std::uint32_t(grpLead) << 3)
                 << (4 * r);
        const std::uint64_t bo = __ballot(own), bl = __ballot(lead), bg = __ballot(grp);
        if (lane == 0)
        {
            waveCnt[r][wave][0] = std::uint32_t(__popcll(bo) + __popcll(bl));
            waveCnt[r][wave][1] = std::uint32_t(__popcll(bg));
        }
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        // exclusive prefix over (row, wave), in place
        std::uint32_t a = 0, b = 0;
        for (int r = 0; r < ROWS; ++r)
            for (int w = 0; w < kThreads / 64; ++w)
            {
                const std::uint32_t ca = waveCnt[r][w][0], cb = waveCnt[r][w][1];
                waveCnt[r][w][0] = a;
                waveCnt[r][w][1] = b;
                a += ca;
                b += cb;
            }
        blockBase[0] = a ? atomicAdd(&t.nfCount[0], a) : 0u;
        blockBase[1] = b ? atomicAdd(&t.nfCount[1], b) : 0u;
    }
    __syncthreads();
    const std::uint64_t lt = (std::uint64_t(1) << lane) - 1;
#pragma unroll 1
    for (int r = 0; r < ROWS; ++r)
    {
        const int i = int((block * unsigned(ROWS) + unsigned(r)) * blockDim.x + threadIdx.x);
        const bool own = (flags >> (4 * r)) & 1u, lead = (flags >> (4 * r + 1)) & 1u,
                   grp = (flags >> (4 * r + 2)) & 1u, grpLead = (flags >> (4 * r + 3)) & 1u;
        const std::uint64_t bo = __ballot(own), bl = __ballot(lead), bg = __ballot(grp);
        const std::uint32_t baseIdx = blockBase[0] + waveCnt[r][wave][0];
        const std::uint32_t baseGrp = blockBase[1] + waveCnt[r][wave][1];
        if (own) t.nfTasks[baseIdx + __popcll(bo & lt)] = std::uint32_t(i) << 1;
        if (lead) t.nfTasks[baseIdx + __popcll(bo) + __popcll(bl & lt)] = (std::uint32_t(i) << 1) | 1u;
        if (grp) t.nfGroups[baseGrp + __popcll(bg & lt)] = (std::uint32_t(i) << 1) | (grpLead ? 1u : 0u);
    }
}
constexpr int kCollectRows = 8; // rows of kThreads leaves per block (single trees)
__global__ __launch_bounds__(kThreads) void nfCollectKernel(TreeDev t, int full)
{
    nfCollectBody<kCollectRows>(t, full, blockIdx.x);
}

// ---- near-field stencils ---------------------------------------------------------
// Cell populations are lower-bound differences over P (computeNFCooperative,
// octree_rebalance.cuh:579-634: the 27 same-level cells around a node,
// periodic at 2^21 per axis). A level-l cell lies inside one level-(l-1)
// cell, whose keys form one contiguous run of P: one global search pair per
// parent-level cell brackets every child-level cell inside it, whose own
// searches then run inside that (short, cache-near) run. The global searches
// start in a table of every S-th entry of the sampled index Q held in LDS by
// each block, then refine through one window of Q and one of P.
constexpr unsigned kTopEntries = 2048;
constexpr int kAxis = int(1u << kOctLevels);

struct TopTable
{
    const Key *keys;
    std::uint32_t count, stride;
};

// The Hilbert tables (hilbert_table::kTables) staged in LDS; every thread
// of the block calls it (before loadTopTable's barrier).
struct HilbertLds
{
    std::uint16_t enc[48 * 8], dec[48 * 8];
};
__device__ void stageHilbert(HilbertLds &h)
{
    for (int i = threadIdx.x; i < 48 * 8; i += blockDim.x)
    {
        h.enc[i] = hilbert_table::kTables.enc[i];
        h.dec[i] = hilbert_table::kTables.dec[i];
    }
}

// Called by every thread of the block.
__device__ TopTable loadTopTable(Key *lds, const Key *Q, std::size_t nq)
{
    const std::uint32_t stride = nq <= kTopEntries ? 1u : std::uint32_t((nq + kTopEntries - 1) / kTopEntries);
    const std::uint32_t count = std::uint32_t((nq + stride - 1) / stride);
    for (std::uint32_t i = threadIdx.x; i < count; i += blockDim.x) lds[i] = Q[std::size_t(i) * stride];
    __syncthreads();
    return TopTable{lds, count, stride};
}

// lower_bound(P, n, v0), lower_bound(P, n, v1) through the LDS table and Q.
__device__ inline void lowerBoundTopPair(const Key *P, std::size_t n, const Key *Q, const TopTable &tab, Key v0,
                                         Key v1, std::size_t &r0, std::size_t &r1)
{
    std::size_t k0 = 0, k1 = 0;
    std::uint32_t m = tab.count;
    if (m)
    {
        std::uint32_t b0 = 0, b1 = 0;
        while (m > 1)
        {
            const std::uint32_t h = m >> 1;
            b0 = tab.keys[b0 + h] < v0 ? b0 + h : b0;
            b1 = tab.keys[b1 + h] < v1 ? b1 + h : b1;
            m -= h;
        }
        k0 = b0 + (tab.keys[b0] < v0 ? 1u : 0u);
        k1 = b1 + (tab.keys[b1] < v1 ? 1u : 0u);
    }
    const std::size_t nq = indexSize(n);
    std::size_t j0, j1;
    refinePair(Q, nq, tab.stride, k0, k1, v0, v1, j0, j1);
    refinePair(P, n, kIndexStride, j0, j1, v0, v1, r0, r1);
}

// The same with 32-bit positions (trees hold fewer than 2^32 particles).
__device__ inline void lowerBoundTopPair32(const Key *P, std::uint32_t n, const Key *Q, const TopTable &tab, Key v0,
                                           Key v1, std::uint32_t &r0, std::uint32_t &r1)
{
    std::uint32_t k0 = 0, k1 = 0;
    std::uint32_t m = tab.count;
    if (m)
    {
        std::uint32_t b0 = 0, b1 = 0;
        while (m > 1)
        {
            const std::uint32_t h = m >> 1;
            b0 = tab.keys[b0 + h] < v0 ? b0 + h : b0;
            b1 = tab.keys[b1 + h] < v1 ? b1 + h : b1;
            m -= h;
        }
        k0 = b0 + (tab.keys[b0] < v0 ? 1u : 0u);
        k1 = b1 + (tab.keys[b1] < v1 ? 1u : 0u);
    }
    const std::uint32_t nq = std::uint32_t(indexSize(n));
    std::uint32_t j0, j1;
    refinePair32(Q, nq, tab.stride, k0, k1, v0, v1, j0, j1);
    refinePair32(P, n, std::uint32_t(kIndexStride), j0, j1, v0, v1, r0, r1);
}

__device__ inline int wrapAxis(int v) { return v < 0 ? v + kAxis : (v >= kAxis ? v - kAxis : v); }

// Near-field count of the level-`level` node at `start` (its own population
// and the sum over its 27 cells), by groups of 32 lanes: lanes 0..7 search
// the 2 x 2 x 2 level-(level-1) cells that hold the stencil, lanes 0..26 then
// count one stencil cell each inside its parent-level run (a cell's key is
// its parent's code plus one digit). The root's stencil (27 copies of the
// root) is searched directly.
__device__ inline void nearField(const Key *P, std::size_t n, const Key *Q, const TopTable &tab,
                                 const std::uint16_t *enc, const std::uint16_t *dec, Key start, unsigned level,
                                 bool hilbert, unsigned lane32, std::uint64_t &center, std::uint64_t &sum)
{
    std::uint64_t pop = 0;
    const std::uint32_t n32 = std::uint32_t(n);
    if (level == 0)
    {
        if (lane32 < 27)
        {
            std::uint32_t r0, r1;
            lowerBoundTopPair32(P, n32, Q, tab, 0, nodeRange(0), r0, r1);
            pop = std::uint64_t(r1 - r0);
        }
    }
    else
    {
        unsigned cx, cy, cz, st;
        cellCorner(start, level, hilbert, dec, cx, cy, cz, st);
        const int w = int(1u << (kOctLevels - level));
        const int u[3] = {int(cx) / w, int(cy) / w, int(cz) / w}; // node index along each axis
        std::uint32_t pa = 0, pb = 0;
        CellCode pc{0, 0};
        if (lane32 < 8)
        {
            int q[3];
            for (int a = 0; a < 3; ++a) q[a] = wrapAxis((((u[a] - 1) >> 1) + int((lane32 >> a) & 1u)) * 2 * w);
            pc = cellCode(unsigned(q[0]), unsigned(q[1]), unsigned(q[2]), level - 1, hilbert, enc);
            lowerBoundTopPair32(P, n32, Q, tab, pc.key, pc.key + nodeRange(level - 1), pa, pb);
        }
        int parent = 0;
        unsigned raw = 0;
        if (lane32 < 27)
        {
            const int d[3] = {int(lane32 % 3) - 1, int((lane32 / 3) % 3) - 1, int(lane32 / 9) - 1};
            for (int a = 0; a < 3; ++a)
            {
                const int v = u[a] + d[a];
                parent |= ((v >> 1) - ((u[a] - 1) >> 1)) << a;
                raw |= unsigned(wrapAxis(v * w) / w & 1) << (2 - a);
            }
        }
        const std::uint32_t lo = __shfl(pa, parent, 32), hi = __shfl(pb, parent, 32);
        const CellCode home{__shfl(pc.key, parent, 32), unsigned(__shfl(int(pc.state), parent, 32))};
        if (lane32 < 27)
        {
            const Key ck = childKey(home, raw, level, hilbert, enc);
            std::uint32_t r0, r1;
            lowerBoundPairIn32(P, lo, hi, ck, ck + nodeRange(level), r0, r1);
            pop = std::uint64_t(r1 - r0);
        }
    }
    std::uint64_t s = pop;
#pragma unroll
    for (int m = 16; m >= 1; m >>= 1) s += __shfl_xor(s, m, 32);
    sum = s;
    center = __shfl(pop, 13, 32);
}

// Two complete sibling groups (level l >= 1) per wave, one per half-wave:
// lanes h*32 + 0..26 name the 27 level-(l-1) cells around group h's parent
// (the parent's own stencil) and search them. The second group's cells that
// are also the first group's (consecutive groups lie close in key order)
// take the first group's bounds instead of searching again. Then, for each
// group in turn, lane c counts the level-l cell (c%4 - 1, c/4%4 - 1,
// c/16 - 1) around the parent's corner (the 4x4x4 block holding the 27-cell
// stencils of all 8 children) inside its parent-level run, and lanes 0..7
// sum their child's 27 cells. Items flagged for it also get the parent's
// merge test (its NF from the same 27 parent-level searches). Same cells and
// the same u64 products as the per-node stencil.
__device__ void nfGroupBody(const TreeDev &t, int hilbert, unsigned bucket, unsigned block, unsigned blocks)
{
    __shared__ Key topKeys[kTopEntries];
    __shared__ HilbertLds hl;
    DevState *st = t.st;
    if (st->converged) return;
    const std::uint32_t total = t.nfCount[1];
    const std::uint32_t wavesPerBlock = blockDim.x / 64;
    if (2 * block * wavesPerBlock >= total) return; // no item for this block
    if (hilbert) stageHilbert(hl);
    const TopTable tab = loadTopTable(topKeys, t.Pidx, indexSize(t.n));
    const int ab = st->activeBuffer;
    const Key *tree = t.K[ab];
    const unsigned lane = __lane_id();
    const unsigned half = lane >> 5, hlane = lane & 31u;
    const std::uint32_t waves = blocks * wavesPerBlock;
    const bool hil = hilbert != 0;
    const std::uint32_t n32 = std::uint32_t(t.n);
    for (std::uint32_t g2 = 2 * (block * wavesPerBlock + (threadIdx.x >> 6)); g2 < total; g2 += 2 * waves)
    {
        // group `half` of the pair (the second half repeats the first when the pair has one group)
        const bool hasB = g2 + 1 < total;
        const std::uint32_t item = t.nfGroups[half && hasB ? g2 + 1 : g2];
        const int leaf0 = int(item >> 1);
        const Key start = tree[leaf0];
        const unsigned level = treeLevel(tree[leaf0 + 1] - start);
        const int w = int(1u << (kOctLevels - level));
        // the parent's corner and its 27 parent-level cells
        unsigned bxu, byu, bzu, pst;
        cellCorner(start, level - 1, hil, hl.dec, bxu, byu, bzu, pst);
        const int bx = int(bxu), by = int(byu), bz = int(bzu);
        const int dx = int(hlane % 3) - 1, dy = int((hlane / 3) % 3) - 1, dz = int(hlane / 9) - 1;
        CellCode pc{0, 0};
        if (hlane < 27)
            pc = cellCode(unsigned(wrapAxis(bx + 2 * dx * w)), unsigned(wrapAxis(by + 2 * dy * w)),
                          unsigned(wrapAxis(bz + 2 * dz * w)), level - 1, hil, hl.enc);
        // a cell of the second group that is also one of the first group's:
        // its lane in the first half (same level; per axis the offset from
        // the first parent, over the level-(l-1) grid of G cells per axis)
        int from = -1;
        const unsigned levelA = __shfl(level, 0, 64);
        const int axA = __shfl(bx, 0, 64), ayA = __shfl(by, 0, 64), azA = __shfl(bz, 0, 64);
        if (half && hasB && hlane < 27 && level == levelA)
        {
            const int G = kAxis / (2 * w);
            auto near = [&](int cB, int cA, int d) {
                // the offset e in -1..1 with (cA/2w + e) = (cB/2w + d) mod G, or 2 if none
                const int target = ((cB / (2 * w) + d) % G + G) % G;
                for (int e = -1; e <= 1; ++e)
                    if ((((cA / (2 * w) + e) % G + G) % G) == target) return e;
                return 2;
            };
            const int ex = near(bx, axA, dx), ey = near(by, ayA, dy), ez = near(bz, azA, dz);
            if (ex != 2 && ey != 2 && ez != 2) from = (ex + 1) + 3 * (ey + 1) + 9 * (ez + 1);
        }
        std::uint32_t pa = 0, pb = 0;
        if (hlane < 27 && from < 0 && (half == 0 || hasB))
            lowerBoundTopPair32(t.P, n32, t.Pidx, tab, pc.key, pc.key + nodeRange(level - 1), pa, pb);
        {
            const int src = from >= 0 ? from : int(lane);
            const std::uint32_t qa = __shfl(pa, src, 64), qb = __shfl(pb, src, 64);
            pa = qa;
            pb = qb;
        }
        // each group in turn: the child-level cells, each inside its
        // parent-level cell's run
        const int a = int(lane & 3u) - 1, b = int((lane >> 2) & 3u) - 1, c = int(lane >> 4) - 1;
        auto third = [](int o) { return o < 0 ? 0 : (o < 2 ? 1 : 2); };
        // the child's octant inside its parent: the coordinate bit at level l
        const unsigned raw = unsigned((a + 2) & 1) << 2 | unsigned((b + 2) & 1) << 1 | unsigned((c + 2) & 1);
        for (unsigned h = 0; h < (hasB ? 2u : 1u); ++h)
        {
            const int base = int(32 * h);
            const std::uint32_t itemH = __shfl(item, base, 64);
            const int leaf0H = int(itemH >> 1);
            const unsigned levelH = __shfl(level, base, 64), pstH = __shfl(pst, base, 64);
            const int parent = base + third(a) + 3 * third(b) + 9 * third(c);
            const std::uint32_t lo = __shfl(pa, parent, 64), hi = __shfl(pb, parent, 64);
            const CellCode home{__shfl(pc.key, parent, 64), unsigned(__shfl(int(pc.state), parent, 64))};
            const Key ck = childKey(home, raw, levelH, hil, hl.enc);
            std::uint32_t r0, r1;
            lowerBoundPairIn32(t.P, lo, hi, ck, ck + nodeRange(levelH), r0, r1);
            const std::uint32_t pop = r1 - r0;
            // child `lane` (lanes 0..7) of the group: its octant inside the
            // parent, from its digit at level l and the parent's state
            const int leaf = leaf0H + int(lane & 7u);
            const unsigned digit = unsigned(tree[leaf] >> (3u * (kOctLevels - levelH))) & 7u;
            const unsigned oct = hil ? unsigned(hl.dec[8 * pstH + digit]) & 7u : digit;
            const int rx = int(oct >> 2), ry = int((oct >> 1) & 1u), rz = int(oct & 1u);
            std::uint64_t sum = 0;
#pragma unroll
            for (int ez = 0; ez < 3; ++ez)
#pragma unroll
                for (int ey = 0; ey < 3; ++ey)
#pragma unroll
                    for (int ex = 0; ex < 3; ++ex)
                        sum += __shfl(pop, (rx + ex) + 4 * (ry + ey) + 16 * (rz + ez), 64);
            const std::uint64_t own = __shfl(pop, (rx + 1) + 4 * (ry + 1) + 16 * (rz + 1), 64);
            if (lane < 8)
            {
                const std::uint64_t nf = own * sum;
                t.N[ab][leaf] = own > 0xffffffffull ? 0xffffffffu : unsigned(own);
                t.NF[ab][leaf] = nf > 0xffffffffull ? 0xffffffffu : unsigned(nf);
            }
            if (itemH & 1u)
            {
                // the parent's merge test: its population times its 27 cells'
                // (groupCanMerge, compared unclamped)
                std::uint64_t pp = half == h && hlane < 27 ? std::uint64_t(pb - pa) : 0;
                const std::uint64_t pown = __shfl(pp, base + 13, 64);
#pragma unroll
                for (int m = 32; m >= 1; m >>= 1) pp += __shfl_xor(pp, m, 64);
                if (lane == 0) t.gcm[ab][leaf0H] = pown * pp <= std::uint64_t(bucket) ? 1 : 0;
            }
        }
    }
}
__global__ __launch_bounds__(kThreads) void nfGroupKernel(TreeDev t, int hilbert, unsigned bucket)
{
    nfGroupBody(t, hilbert, bucket, blockIdx.x, gridDim.x);
}

__device__ void nfComputeBody(const TreeDev &t, unsigned bucket, int hilbert, unsigned block, unsigned blocks)
{
    __shared__ Key topKeys[kTopEntries];
    __shared__ HilbertLds hl;
    DevState *st = t.st;
    if (st->converged) return;
    const std::uint32_t total = t.nfCount[0];
    if (block * (blockDim.x / 32) >= total) return; // no item for this block
    if (hilbert) stageHilbert(hl);
    const TopTable tab = loadTopTable(topKeys, t.Pidx, indexSize(t.n));
    const int ab = st->activeBuffer;
    const Key *tree = t.K[ab];
    const unsigned lane32 = threadIdx.x & 31;
    const std::uint32_t groups = blocks * (blockDim.x / 32);
    for (std::uint32_t g = block * (blockDim.x / 32) + threadIdx.x / 32; g < total; g += groups)
    {
        const std::uint32_t task = t.nfTasks[g];
        const int leaf = int(task >> 1);
        const bool parent = task & 1u;
        const Key start = tree[leaf];
        unsigned level = treeLevel(tree[leaf + 1] - start);
        if (parent) level -= 1;
        std::uint64_t center, sum;
        nearField(t.P, t.n, t.Pidx, tab, hl.enc, hl.dec, start, level, hilbert != 0, lane32, center, sum);
        if (lane32 == 0)
        {
            const std::uint64_t nf = center * sum;
            if (!parent)
            {
                t.N[ab][leaf] = center > 0xffffffffull ? 0xffffffffu : unsigned(center);
                t.NF[ab][leaf] = nf > 0xffffffffull ? 0xffffffffu : unsigned(nf);
            }
            else
                t.gcm[ab][leaf] = nf <= std::uint64_t(bucket) ? 1 : 0;
        }
    }
}
__global__ __launch_bounds__(kThreads) void nfComputeKernel(TreeDev t, unsigned bucket, int hilbert)
{
    nfComputeBody(t, bucket, hilbert, blockIdx.x, gridDim.x);
}

// ---- one-pass converged binary tree ---------------------------------------------
// Internal nodes are exactly the nodes that contain a window of bucket+1
// consecutive sorted particles (count > bucket) above depth maxDepth. With
// D_j = min(cpl(P[j], P[j+bucket]), maxDepth-1) for each window j, the
// deepest internal node holding particle q sits at depth
//   delta(q) = max D_j over the windows j in [q-bucket, q].
// The leaf boundaries are {0} and the midpoints of the internal nodes; those
// lying in (P[q-1], P[q]] are, in ascending order,
//   (a) internal ancestors of P[q-1] below the common ancestor where P[q-1]
//       lies in the lower half, deepest first,
//   (b) the internal common ancestor of P[q-1] and P[q],
//   (c) internal ancestors of P[q] below the common ancestor where P[q] lies
//       in the upper half, shallowest first,
// and every boundary written for q has lower_bound(P, boundary) == q.
__device__ inline int cpl63(Key a, Key b) { return clz64(a ^ b) - 1; } // 63 when equal

__device__ inline Key bitsBetween(int lo, int hi) // bits lo..hi inclusive, 0 <= lo, hi <= 62
{
    if (hi < lo) return 0;
    const Key top = (Key(1) << (hi + 1)) - 1;
    return top & ~((Key(1) << lo) - 1);
}

constexpr int kQ = kPassTile; // particles per tile of the binary construction
constexpr int kWindowBucket = 64; // buckets up to this take the register windows, larger a sparse table

// Per tile of kQ particles: delta(q) via a sparse table over the D window in
// LDS, the number of boundaries of each interval q, and the tile's total.
__global__ __launch_bounds__(kThreads) void binCountKernel(const Key *P, const DevState *st, unsigned bucket,
                                                           unsigned maxDepth, std::int8_t *delta, std::uint8_t *cnt,
                                                           std::uint32_t *tileSum)
{
    // dynamic LDS sized for this bucket (binCountLdsBytes): keys P[q0-1-B ..
    // q0+kQ+B], two rows of the sparse table, delta per interval
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ std::uint32_t part[kThreads / 64];
    const int tid = threadIdx.x;
    const long nact = long(st->numActiveParticles);
    const long q0 = long(blockIdx.x) * kQ;
    if (q0 > nact) return;
    const int B = int(bucket);
    const long origin = q0 - 1 - B; // particle index of pw[0] and tab[.][0]
    const int spanP = kQ + 2 * B + 2;
    const int spanD = kQ + B + 1;
    const int rowD = (spanD + 7) & ~7; // 8-byte aligned rows
    Key *pw = reinterpret_cast<Key *>(lds);
    std::int8_t *tab[2] = {reinterpret_cast<std::int8_t *>(pw + spanP),
                           reinterpret_cast<std::int8_t *>(pw + spanP) + rowD};
    std::int8_t *dl = tab[1] + rowD;
    // eight loads in flight per thread before the first LDS store
    for (int s0 = 0; s0 < spanP; s0 += 8 * kThreads)
    {
        Key v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
        {
            const int s = s0 + k * kThreads + tid;
            const long j = origin + s;
            v[k] = (s < spanP && j >= 0 && j < nact) ? P[j] : Key(0);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
        {
            const int s = s0 + k * kThreads + tid;
            if (s < spanP) pw[s] = v[k];
        }
    }
    __syncthreads();
    for (int s = tid; s < spanD; s += kThreads)
    {
        const long j = origin + s;
        int D = -1;
        if (maxDepth > 0 && j >= 0 && j + B < nact)
        {
            const int c = cpl63(pw[s], pw[s + B]);
            D = c < int(maxDepth) - 1 ? c : int(maxDepth) - 1;
        }
        tab[0][s] = std::int8_t(D);
    }
    __syncthreads();
    // dl[u] = delta(q0 - 1 + u): windows [q-B, q] are tab indices [u, u+B]
    if (B <= kWindowBucket)
    {
        // Thread t: the nine windows u = 8t .. 8t+8 in registers. They share
        // tab[8t+8 .. 8t+B]; window i adds a suffix of tab[8t .. 8t+7] on the
        // left and a prefix of tab[8t+B+1 .. 8t+B+8] on the right.
        static_assert(kItems == 8, "eight windows per thread");
        const int u0 = tid * kItems;
        const std::int8_t *t0 = tab[0];
        int win[kItems + 1];
        if (B >= kItems)
        {
            int common = -128;
            for (int o = kItems; o <= B; ++o) common = max(common, int(t0[u0 + o]));
            int sl[kItems + 1];
            sl[kItems] = -128;
#pragma unroll
            for (int i = kItems - 1; i >= 0; --i) sl[i] = max(sl[i + 1], int(t0[u0 + i]));
            int pr = -128;
            win[0] = max(sl[0], common);
#pragma unroll
            for (int i = 1; i <= kItems; ++i)
            {
                pr = max(pr, int(t0[u0 + B + i]));
                win[i] = max(max(sl[i], common), pr);
            }
        }
        else
        {
#pragma unroll
            for (int i = 0; i <= kItems; ++i)
            {
                int w = -128;
                for (int o = 0; o <= B; ++o) w = max(w, int(t0[u0 + i + o]));
                win[i] = w;
            }
        }
        std::uint64_t packed = 0;
#pragma unroll
        for (int i = 0; i <= kItems; ++i)
        {
            const long q = q0 - 1 + u0 + i;
            const int d = (q >= 0 && q < nact) ? win[i] : -1;
            if (i < kItems) packed |= std::uint64_t(std::uint8_t(d)) << (8 * i);
            else if (tid == kThreads - 1) dl[kQ] = std::int8_t(d);
        }
        *reinterpret_cast<std::uint64_t *>(dl + u0) = packed;
    }
    else
    {
        // sparse table: after round t, tab[cur][s] = max of tab[0][s .. s + 2^t - 1]
        int lg = 0;
        while ((2 << lg) <= B + 1) ++lg; // floor(log2(bucket + 1))
        int cur = 0;
        for (int t = 0; t < lg; ++t)
        {
            const int step = 1 << t;
            for (int s = tid; s < spanD; s += kThreads)
            {
                std::int8_t v = tab[cur][s];
                if (s + step < spanD)
                {
                    const std::int8_t w = tab[cur][s + step];
                    v = w > v ? w : v;
                }
                tab[cur ^ 1][s] = v;
            }
            __syncthreads();
            cur ^= 1;
        }
        const int pw2 = 1 << lg;
        for (int u = tid; u <= kQ; u += kThreads)
        {
            const long q = q0 - 1 + u;
            int d = -1;
            if (q >= 0 && q < nact)
            {
                const int a = tab[cur][u], b = tab[cur][u + B - pw2 + 1];
                d = a > b ? a : b;
            }
            dl[u] = std::int8_t(d);
        }
    }
    __syncthreads();
    std::uint32_t sum = 0;
    for (int u = tid; u < kQ; u += kThreads)
    {
        const long q = q0 + u;
        if (q > nact) break;
        const bool hp = q > 0, hc = q < nact;
        const Key prev = pw[u + B], curk = pw[u + B + 1]; // P[q-1], P[q]
        const int dp = hp ? dl[u] : -1;
        const int dc = hc ? dl[u + 1] : -1;
        const int cp = (hp && hc) ? cpl63(prev, curk) : -1;
        const int lca = (hp && hc && cp <= dc) ? 1 : 0;
        const Key am = dp > cp ? bitsBetween(62 - dp, 61 - cp) : 0;
        const Key bm = dc > cp ? bitsBetween(62 - dc, 61 - cp) : 0;
        const int c = __popcll(~prev & am) + lca + __popcll(curk & bm);
        cnt[q] = std::uint8_t(c);
        if (hc) delta[q] = std::int8_t(dc);
        sum += std::uint32_t(c);
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) sum += __shfl_xor(sum, m, 64);
    if (__lane_id() == 0) part[tid >> 6] = sum;
    __syncthreads();
    if (tid == 0)
    {
        std::uint32_t s = 0;
        for (int w = 0; w < kThreads / 64; ++w) s += part[w];
        tileSum[blockIdx.x] = s;
    }
}

// Exclusive scan of the tile sums (one block) and the total boundary count.
__global__ __launch_bounds__(1024) void binScanTilesKernel(const DevState *st, std::uint32_t *tileSum,
                                                          unsigned long long *total)
{
    __shared__ unsigned long long wsum[16];
    const long tiles = long(st->numActiveParticles) / kQ + 1;
    const int tid = threadIdx.x;
    const long per = (tiles + 1023) / 1024;
    const long b = long(tid) * per, e = b + per < tiles ? b + per : tiles;
    constexpr int kBatch = 8; // loads in flight per thread
    unsigned long long local = 0;
    for (long i0 = b; i0 < e; i0 += kBatch)
    {
        std::uint32_t v[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; ++k) v[k] = i0 + k < e ? tileSum[i0 + k] : 0u;
#pragma unroll
        for (int k = 0; k < kBatch; ++k) local += v[k];
    }
    const unsigned lane = __lane_id();
    const int wave = tid >> 6;
    unsigned long long inc = local;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const unsigned long long t = __shfl_up(inc, o, 64);
        if (lane >= unsigned(o)) inc += t;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    unsigned long long run = inc - local;
    for (int w = 0; w < wave; ++w) run += wsum[w];
    for (long i0 = b; i0 < e; i0 += kBatch)
    {
        std::uint32_t v[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; ++k) v[k] = i0 + k < e ? tileSum[i0 + k] : 0u;
#pragma unroll
        for (int k = 0; k < kBatch; ++k)
            if (i0 + k < e)
            {
                tileSum[i0 + k] = std::uint32_t(run);
                run += v[k];
            }
    }
    if (tid == 1023) *total = run;
}

// Writes the boundaries of each interval q at 1 + (boundaries before q).
// Few intervals hold boundaries, so each wave first lists its non-empty
// intervals (with their output positions) in LDS and then emits them one
// per lane: the loads of P and delta for a wave's intervals go out together.
// The leaf count is total + 1; nothing is written when it exceeds capacity
// (the host grows the buffers and runs the emission again).
__global__ __launch_bounds__(kThreads) void binEmitKernel(const Key *P, const DevState *st, const std::int8_t *delta,
                                                          const std::uint8_t *cnt, const std::uint32_t *tileOff,
                                                          Key *K, std::uint32_t *bound,
                                                          const unsigned long long *total, long capacity)
{
    constexpr int kW = kThreads / 64;
    __shared__ std::uint32_t wsum[kW];
    __shared__ uint2 list[kW][64 * kItems]; // (interval, first output position)
    const long leaves = long(*total) + 1;
    if (leaves > capacity) return;
    const int tid = threadIdx.x;
    const long nact = long(st->numActiveParticles);
    const long q0 = long(blockIdx.x) * kQ;
    if (q0 > nact) return;
    static_assert(kItems == 8, "one 8-byte load of interval counts per thread");
    const long qFirst = q0 + long(tid) * kItems;
    const std::uint32_t tileBase = tileOff[blockIdx.x];
    // cnt is allocated to whole tiles; entries past nact are masked
    const uint2 raw = *reinterpret_cast<const uint2 *>(cnt + qFirst);
    std::uint32_t c[kItems];
    std::uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        const long q = qFirst + k;
        const std::uint32_t v = ((k < 4 ? raw.x : raw.y) >> (8 * (k & 3))) & 0xffu;
        c[k] = q <= nact ? v : 0u;
        sum += c[k];
    }
    const unsigned lane = __lane_id();
    const int wave = tid >> 6;
    std::uint32_t inc = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const std::uint32_t t = __shfl_up(inc, o, 64);
        if (lane >= unsigned(o)) inc += t;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    std::uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += wsum[w];
    long pos = 1 + long(tileBase) + long(before + inc - sum);
    const std::uint64_t ltMask = (std::uint64_t(1) << lane) - 1;
    std::uint32_t listed = 0; // wave-uniform
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        const long q = qFirst + k;
        const bool has = c[k] != 0;
        const std::uint64_t b = __ballot(has);
        if (has) list[wave][listed + __popcll(b & ltMask)] = make_uint2(std::uint32_t(q), std::uint32_t(pos));
        listed += std::uint32_t(__popcll(b));
        pos += long(c[k]);
        if (q == nact)
        {
            K[0] = 0;
            bound[0] = 0;
            if (pos == leaves)
            {
                K[leaves] = kMaxKey;
                bound[leaves] = std::uint32_t(nact);
            }
        }
    }
    __syncthreads();
    for (std::uint32_t e = lane; e < listed; e += 64)
    {
        const uint2 it = list[wave][e];
        const long q = long(it.x);
        long at = long(it.y);
        const bool hp = q > 0, hc = q < nact;
        const Key prev = hp ? P[q - 1] : 0, curk = hc ? P[q] : 0;
        const int dp = hp ? delta[q - 1] : -1;
        const int dc = hc ? delta[q] : -1;
        const int cp = (hp && hc) ? cpl63(prev, curk) : -1;
        Key m = (dp > cp ? bitsBetween(62 - dp, 61 - cp) : 0) & ~prev;
        while (m)
        {
            const int bit = __ffsll((long long)m) - 1;
            const Key mid = (prev & ~((Key(2) << bit) - 1)) | (Key(1) << bit);
            if (at < leaves) { K[at] = mid; bound[at] = std::uint32_t(q); }
            ++at;
            m &= m - 1;
        }
        if (hp && hc && cp <= dc)
        {
            const int bit = 62 - cp;
            if (at < leaves) { K[at] = curk & ~((Key(1) << bit) - 1); bound[at] = std::uint32_t(q); }
            ++at;
        }
        m = (dc > cp ? bitsBetween(62 - dc, 61 - cp) : 0) & curk;
        while (m)
        {
            const int bit = 63 - clz64(m);
            if (at < leaves) { K[at] = curk & ~((Key(1) << bit) - 1); bound[at] = std::uint32_t(q); }
            ++at;
            m &= ~(Key(1) << bit);
        }
    }
}

__global__ __launch_bounds__(kThreads) void binBoundsToCountsKernel(const std::uint32_t *bound, std::uint32_t *N,
                                                                    const unsigned long long *total, long capacity)
{
    const long leaves = long(*total) + 1;
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (leaves <= capacity && i < leaves) N[i] = bound[i + 1] - bound[i];
}

// The state of the constructed tree (its leaf count also when it did not
// fit the capacity: the host then grows and emits again).
__global__ void binDirectStateKernel(DevState *st, const unsigned long long *total)
{
    const int leaves = int(*total + 1);
    st->changed = 0;
    st->converged = 1;
    st->numLeaves = leaves;
    st->newNumLeaves = leaves;
    st->needsResize = 0;
    st->numNodes = 2 * leaves - 1;
    st->activeBuffer = 0;
}

__device__ void buildIndexBody(const Key *P, std::size_t n, Key *Q, unsigned block, unsigned blocks)
{
    const std::size_t nq = indexSize(n);
    for (std::size_t j = std::size_t(block) * blockDim.x + threadIdx.x; j < nq; j += std::size_t(blocks) * blockDim.x)
        Q[j] = P[j * kIndexStride];
}
__global__ void buildIndexKernel(const Key *P, std::size_t n, Key *Q) { buildIndexBody(P, n, Q, blockIdx.x, gridDim.x); }

unsigned clearBlocks(std::size_t words)
{
    const unsigned b = ceilDiv(words, 256);
    return b < 1 ? 1 : (b > 1024 ? 1024 : b);
}

// Tile b (kBoundTile keys of P) owns the leaf boundaries i with
// P[b * kBoundTile - 1] < K[i] <= P[(b + 1) * kBoundTile - 1]; tileFirst[b]
// is its first. Entries b = block * kThreads + lane, striding by blocks.
__device__ void tileFirstBody(const TreeDev &t, unsigned block, unsigned blocks)
{
    const DevState *st = t.st;
    if (st->converged || t.n == 0) return;
    const std::size_t tiles = (t.n + kBoundTile - 1) / kBoundTile;
    const std::uint32_t L = std::uint32_t(st->numLeaves);
    const Key *K = t.K[st->activeBuffer];
    for (std::size_t b = std::size_t(block) * kThreads + threadIdx.x; b <= tiles; b += std::size_t(blocks) * kThreads)
        t.tileFirst[b] = b == 0 ? 0u
                         : b == tiles ? L + 1
                                      : std::uint32_t(upperBound(K, std::size_t(L) + 1, t.P[b * kBoundTile - 1]));
}
__global__ __launch_bounds__(kThreads) void tileFirstKernel(TreeDev t) { tileFirstBody(t, blockIdx.x, gridDim.x); }
// Tile `tile`: its keys in LDS, then the lower bound of each boundary it owns
// (into t.offs, read as 32-bit words).
__device__ void tileBoundsBody(const TreeDev &t, unsigned tile)
{
    __shared__ Key tp[kBoundTile];
    const DevState *st = t.st;
    if (st->converged) return;
    const Key *K = t.K[st->activeBuffer];
    const std::size_t s = std::size_t(tile) * kBoundTile;
    const std::size_t len = t.n - s < kBoundTile ? t.n - s : kBoundTile;
    {
        constexpr int kPer = kBoundTile / kThreads;
        Key v[kPer]; // all loads in flight before the first LDS store
#pragma unroll
        for (int k = 0; k < kPer; ++k)
        {
            const std::size_t i = threadIdx.x + std::size_t(k) * kThreads;
            v[k] = i < len ? t.P[s + i] : Key(0);
        }
#pragma unroll
        for (int k = 0; k < kPer; ++k)
        {
            const std::size_t i = threadIdx.x + std::size_t(k) * kThreads;
            if (i < len) tp[i] = v[k];
        }
    }
    __syncthreads();
    std::uint32_t *lb = reinterpret_cast<std::uint32_t *>(t.offs);
    const std::uint32_t i1 = t.tileFirst[tile + 1];
    for (std::uint32_t i = t.tileFirst[tile] + threadIdx.x; i < i1; i += kThreads)
        lb[i] = std::uint32_t(s + lowerBound(tp, len, K[i]));
}
__global__ __launch_bounds__(kThreads) void tileBoundsKernel(TreeDev t) { tileBoundsBody(t, blockIdx.x); }
// Leaves i = block * kThreads + lane, striding by blocks: N[i] from the bounds.
__device__ void tileCountsBody(const TreeDev &t, unsigned block, unsigned blocks)
{
    const DevState *st = t.st;
    if (st->converged) return;
    const int L = st->numLeaves;
    const std::uint32_t *lb = reinterpret_cast<const std::uint32_t *>(t.offs);
    std::uint32_t *N = t.N[st->activeBuffer];
    for (int i = int(block * kThreads + threadIdx.x); i < L; i += int(blocks * kThreads))
        N[i] = t.n == 0 ? 0u : lb[i + 1] - lb[i];
}
__global__ __launch_bounds__(kThreads) void tileCountsKernel(TreeDev t) { tileCountsBody(t, blockIdx.x, gridDim.x); }

// ---- multi-tree kernels: block -> (tree, block within the tree) ------------------
__device__ inline int treeOfBlock(const TreeSet &s, unsigned b)
{
    int lo = 0, hi = s.count - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (s.start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The member's TreeDev in LDS for the block: the bodies index its buffer
// pairs by the active half at run time, which a copy in registers would
// place in per-lane scratch memory. Every thread of the block calls it.
__device__ inline void stageTreeDev(TreeDev &lds, const TreeDev &src)
{
    static_assert(sizeof(TreeDev) % 4 == 0, "TreeDev is copied as 32-bit words");
    const std::uint32_t *from = reinterpret_cast<const std::uint32_t *>(&src);
    std::uint32_t *to = reinterpret_cast<std::uint32_t *>(&lds);
    for (unsigned w = threadIdx.x; w < sizeof(TreeDev) / 4; w += blockDim.x) to[w] = from[w];
    __syncthreads();
}

// Every block walks the virtual blocks b = blockIdx.x, + gridDim.x, ... below
// the set's total (one step when the grid equals a host-planned total): the
// tree k that owns b, its TreeDev t and b's index `block` among the tree's
// `blocks`. b is block-uniform, and each step ends with a barrier, so the
// bodies' shared memory is free again for the next step.
#define AT_MULTI_LOOP(BODY)                                                                                            \
    __shared__ alignas(alignof(TreeDev)) unsigned char multiTreeBytes[sizeof(TreeDev)];                               \
    TreeDev &multiTree = *reinterpret_cast<TreeDev *>(multiTreeBytes);                                                 \
    const unsigned virtualBlocks = s.total ? *s.total : s.blocks;                                                      \
    for (unsigned vb = blockIdx.x; vb < virtualBlocks; vb += gridDim.x)                                                \
    {                                                                                                                  \
        const int k = treeOfBlock(s, vb);                                                                              \
        stageTreeDev(multiTree, s.trees[k]);                                                                           \
        const TreeDev &t = multiTree;                                                                                  \
        const unsigned block = vb - s.start[k];                                                                        \
        const unsigned blocks = s.start[k + 1] - s.start[k];                                                           \
        (void)block;                                                                                                   \
        (void)blocks;                                                                                                  \
        BODY;                                                                                                          \
        __syncthreads();                                                                                               \
    }

__global__ void multiOctInitKernel(TreeSet s) { AT_MULTI_LOOP(octInitBody(t, t.statusWords, block, blocks)); }
__global__ void multiBuildIndexKernel(TreeSet s)
{
    AT_MULTI_LOOP(buildIndexBody(t.P, t.n, const_cast<Key *>(t.Pidx), block, blocks));
}
__global__ void multiResetKernel(TreeSet s)
{
    AT_MULTI_LOOP(resetBody(t.st, t.P, t.n, t.scanStatus, t.statusWords, t.scanTicket, t.nfCount, block, blocks));
}
// Clears the chained-scan status words and ticket of each tree (the build
// loop's per-pass reset, without a memset per tree).
__device__ void clearScanBody(const TreeDev &t, unsigned block, unsigned blocks)
{
    if (block == 0 && threadIdx.x == 0) *t.scanTicket = 0;
    for (std::size_t w = std::size_t(block) * blockDim.x + threadIdx.x; w < t.statusWords;
         w += std::size_t(blocks) * blockDim.x)
        t.scanStatus[w] = 0;
}
__global__ void multiClearScanKernel(TreeSet s) { AT_MULTI_LOOP(clearScanBody(t, block, blocks)); }
__device__ void multiOctDecideBody(const TreeDev &t, unsigned bucket, int nf)
{
    const OctDecider d{{t.K[0], t.K[1]}, {t.N[0], t.N[1]}, {t.NF[0], t.NF[1]}, {t.gcm[0], t.gcm[1]}, bucket, nf != 0};
    decideScanBody<kThreads>(d, t.st, t.capacity, t.offs, t.scanStatus, t.scanTicket);
}
__global__ __launch_bounds__(kThreads) void multiOctDecideScanKernel(TreeSet s, unsigned bucket, int nf)
{
    AT_MULTI_LOOP(multiOctDecideBody(t, bucket, nf));
}
// Leaf-shaped multi-tree launches: a virtual block covers kLeafGroup rows of
// kThreads leaves of one tree (the member lookup and the TreeDev load are
// paid once per 2048 leaves).
constexpr unsigned kLeafGroup = 8;
__device__ void multiProcessGroup(const TreeDev &t, int nf, unsigned block)
{
    for (unsigned r = 0; r < kLeafGroup; ++r) octProcessBody(t, nf != 0, block * kLeafGroup + r);
}
__device__ void multiLeafCountsGroup(const TreeDev &t, unsigned block, int full)
{
    for (unsigned r = 0; r < kLeafGroup; ++r) leafCountsBody(t, block * kLeafGroup + r, full);
}
__global__ __launch_bounds__(kThreads) void multiOctProcessKernel(TreeSet s, int nf)
{
    AT_MULTI_LOOP(multiProcessGroup(t, nf, block));
}
__global__ void multiCommitKernel(TreeSet s, int binary)
{
    AT_MULTI_LOOP(commitBody(t.st, binary, t.scanStatus, t.statusWords, t.scanTicket, t.nfCount, block, blocks,
                             t.passLog));
}
__global__ __launch_bounds__(kThreads) void multiLeafCountsKernel(TreeSet s, int full)
{
    AT_MULTI_LOOP(multiLeafCountsGroup(t, block, full));
}
__global__ __launch_bounds__(kThreads) void multiNfCollectKernel(TreeSet s, int full)
{
    AT_MULTI_LOOP(nfCollectBody<kLeafGroup>(t, full, block));
}
__global__ __launch_bounds__(kThreads) void multiNfComputeKernel(TreeSet s, unsigned bucket, int hilbert)
{
    AT_MULTI_LOOP(nfComputeBody(t, bucket, hilbert, block, blocks));
}
__global__ __launch_bounds__(kThreads) void multiNfGroupKernel(TreeSet s, int hilbert, unsigned bucket)
{
    AT_MULTI_LOOP(nfGroupBody(t, hilbert, bucket, block, blocks));
}
// Full leaf counts by tiles of P (small leaves): kTileFirst blocks of
// kThreads first-boundary searches, one block per tile, then leaf-shaped counts.
__global__ __launch_bounds__(kThreads) void multiTileFirstKernel(TreeSet s)
{
    AT_MULTI_LOOP(tileFirstBody(t, block, blocks));
}
// kTileGroup tiles per virtual block (the member lookup paid once per group)
constexpr unsigned kTileGroup = 8;
__device__ void multiTileBoundsGroup(const TreeDev &t, unsigned block)
{
    const unsigned tiles = unsigned((t.n + kBoundTile - 1) / kBoundTile);
    for (unsigned r = 0; r < kTileGroup; ++r)
    {
        const unsigned tile = block * kTileGroup + r;
        if (tile >= tiles) break; // block-uniform
        tileBoundsBody(t, tile);
        __syncthreads(); // the tile's LDS copy is rewritten by the next one
    }
}
__global__ __launch_bounds__(kThreads) void multiTileBoundsKernel(TreeSet s)
{
    AT_MULTI_LOOP(multiTileBoundsGroup(t, block));
}
__device__ void multiTileCountsGroup(const TreeDev &t, unsigned block, unsigned blocks)
{
    tileCountsBody(t, block, blocks);
}
__global__ __launch_bounds__(kThreads) void multiTileCountsKernel(TreeSet s)
{
    AT_MULTI_LOOP(multiTileCountsGroup(t, block, blocks));
}
#undef AT_MULTI_LOOP

// Device versions of the per-tree block counts (leafBlocksFor, nfBlocksFor,
// tileBlocksFor below).
__device__ inline std::uint32_t dLeafBlocks(std::uint32_t leaves)
{
    constexpr std::uint32_t kLeaves = kThreads * kLeafGroup;
    const std::uint32_t b = (leaves + kLeaves - 1) / kLeaves;
    return b < 1 ? 1u : b;
}
__device__ inline std::uint32_t dNfBlocks(std::uint32_t leaves)
{
    const std::uint32_t b = (leaves + 63) / 64;
    return b < 1 ? 1u : (b > 2048 ? 2048u : b);
}

constexpr int kPlanThreads = 1024;
__global__ __launch_bounds__(kPlanThreads) void multiPlanKernel(const TreeDev *trees, int count, int phase, int nf,
                                                                std::uint32_t *start)
{
    __shared__ std::uint32_t waveSum[kPlanShapes][kPlanThreads / 64];
    __shared__ std::uint32_t carry[kPlanShapes];
    const unsigned lane = __lane_id();
    const int wave = threadIdx.x >> 6;
    if (threadIdx.x < kPlanShapes) carry[threadIdx.x] = 0;
    __syncthreads();
    const std::size_t stride = std::size_t(count) + 1;
    for (int base = 0; base < count; base += kPlanThreads)
    {
        const int k = base + int(threadIdx.x);
        std::uint32_t c[kPlanShapes] = {0u, 0u, 0u, 0u, 0u};
        if (k < count)
        {
            const DevState *st = trees[k].st;
            if (!st->converged)
            {
                const std::uint32_t L = std::uint32_t(st->numLeaves);
                if (phase == 0)
                    c[kPlanTiles] = L / kPassTile + 1;
                else
                {
                    const std::uint32_t newL = std::uint32_t(st->newNumLeaves);
                    c[kPlanProcess] = dLeafBlocks(L);
                    c[kPlanCommit] = 1;
                    c[kPlanLeaves] = dLeafBlocks(newL);
                    c[kPlanNf] = nf ? dNfBlocks(newL) : 0u;
                }
            }
        }
        std::uint32_t inc[kPlanShapes];
#pragma unroll
        for (int sh = 0; sh < kPlanShapes; ++sh)
        {
            inc[sh] = c[sh];
#pragma unroll
            for (int o = 1; o < 64; o <<= 1)
            {
                const std::uint32_t v = __shfl_up(inc[sh], o, 64);
                if (lane >= unsigned(o)) inc[sh] += v;
            }
            if (lane == 63) waveSum[sh][wave] = inc[sh];
        }
        __syncthreads();
#pragma unroll
        for (int sh = 0; sh < kPlanShapes; ++sh)
        {
            std::uint32_t before = carry[sh];
            std::uint32_t all = 0;
            for (int w = 0; w < kPlanThreads / 64; ++w)
            {
                if (w < wave) before += waveSum[sh][w];
                all += waveSum[sh][w];
            }
            const bool mine = phase == 0 ? sh == kPlanTiles : sh != kPlanTiles;
            if (mine && k < count) start[sh * stride + std::size_t(k)] = before + inc[sh] - c[sh];
            __syncthreads(); // every thread has read carry[sh]
            if (threadIdx.x == 0) carry[sh] += all;
        }
        __syncthreads();
    }
    if (threadIdx.x < kPlanShapes)
    {
        const int sh = int(threadIdx.x);
        const bool mine = phase == 0 ? sh == kPlanTiles : sh != kPlanTiles;
        if (mine) start[sh * stride + std::size_t(count)] = carry[sh];
    }
}
} // namespace

unsigned clearBlocksFor(std::size_t statusWords) { return clearBlocks(statusWords); }
unsigned indexBlocksFor(std::size_t n)
{
    const unsigned b = ceilDiv(indexSize(n), 256);
    return b < 1 ? 1 : (b > 1024 ? 1024 : b);
}
void buildSearchIndex(hipStream_t s, const TreeDev &t)
{
    if (t.n) hipLaunchKernelGGL(buildIndexKernel, dim3(indexBlocksFor(t.n)), dim3(256), 0, s, t.P, t.n,
                                const_cast<Key *>(t.Pidx));
}
// Grid of a multi-tree launch: the host-planned total, or a fixed grid that
// loops over a device-planned one (8 blocks of 256 threads per CU).
constexpr unsigned kPlannedGrid = 2048;
inline unsigned multiGrid(const TreeSet &set) { return set.total ? kPlannedGrid : set.blocks; }

void multiBuildSearchIndex(hipStream_t s, const TreeSet &index)
{
    if (multiGrid(index)) hipLaunchKernelGGL(multiBuildIndexKernel, dim3(multiGrid(index)), dim3(256), 0, s, index);
}
unsigned tileBlocksFor(std::size_t gridLeaves) { return unsigned(gridLeaves / kPassTile + 1); }
unsigned leafBlocksFor(std::size_t gridLeaves)
{
    const unsigned b = ceilDiv(gridLeaves, kThreads * kLeafGroup);
    return b < 1 ? 1 : b;
}
unsigned nfBlocksFor(std::size_t gridLeaves)
{
    const unsigned b = ceilDiv(gridLeaves, 64);
    return b < 1 ? 1 : (b > 2048 ? 2048 : b);
}

void multiOctInit(hipStream_t s, const TreeSet &clear)
{
    if (multiGrid(clear)) hipLaunchKernelGGL(multiOctInitKernel, dim3(multiGrid(clear)), dim3(256), 0, s, clear);
}
void multiReset(hipStream_t s, const TreeSet &clear)
{
    if (multiGrid(clear)) hipLaunchKernelGGL(multiResetKernel, dim3(multiGrid(clear)), dim3(256), 0, s, clear);
}
void multiClearScan(hipStream_t s, const TreeSet &clear)
{
    if (multiGrid(clear)) hipLaunchKernelGGL(multiClearScanKernel, dim3(multiGrid(clear)), dim3(256), 0, s, clear);
}
void multiOctDecideScan(hipStream_t s, const TreeSet &tiles, const OctParams &p)
{
    prof::Scope scope("decide_scan", s);
    if (multiGrid(tiles))
        hipLaunchKernelGGL(multiOctDecideScanKernel, dim3(multiGrid(tiles)), dim3(kThreads), 0, s, tiles, p.bucket,
                           p.nf ? 1 : 0);
}
void multiOctProcess(hipStream_t s, const TreeSet &leaves, const OctParams &p)
{
    if (multiGrid(leaves))
        hipLaunchKernelGGL(multiOctProcessKernel, dim3(multiGrid(leaves)), dim3(kThreads), 0, s, leaves,
                           p.nf ? 1 : 0);
}
void multiCommit(hipStream_t s, const TreeSet &clear, bool binary)
{
    if (multiGrid(clear))
        hipLaunchKernelGGL(multiCommitKernel, dim3(multiGrid(clear)), dim3(256), 0, s, clear, binary ? 1 : 0);
}
void multiOctRefresh(hipStream_t s, const TreeSet &leaves, const TreeSet &nf, const OctParams &p, bool full)
{
    if (!multiGrid(leaves)) return;
    if (!p.nf)
    {
        prof::Scope scope("leaf_counts", s);
        hipLaunchKernelGGL(multiLeafCountsKernel, dim3(multiGrid(leaves)), dim3(kThreads), 0, s, leaves,
                           full ? 1 : 0);
        return;
    }
    prof::Scope scope("near_field", s);
    hipLaunchKernelGGL(multiNfCollectKernel, dim3(multiGrid(leaves)), dim3(kThreads), 0, s, leaves, full ? 1 : 0);
    if (multiGrid(nf))
    {
        hipLaunchKernelGGL(multiNfComputeKernel, dim3(multiGrid(nf)), dim3(kThreads), 0, s, nf, p.bucket,
                           p.hilbert ? 1 : 0);
        hipLaunchKernelGGL(multiNfGroupKernel, dim3(multiGrid(nf)), dim3(kThreads), 0, s, nf, p.hilbert ? 1 : 0,
                           p.bucket);
    }
}

unsigned tileFirstBlocksFor(std::size_t n)
{
    return n ? ceilDiv((n + kBoundTile - 1) / kBoundTile + 1, kThreads) : 0u;
}
unsigned tileBlocksOfP(std::size_t n) { return unsigned(((n + kBoundTile - 1) / kBoundTile + kTileGroup - 1) / kTileGroup); }
void multiLeafCountsByTiles(hipStream_t s, const TreeSet &first, const TreeSet &tiles, const TreeSet &leaves)
{
    prof::Scope scope("leaf_counts", s);
    if (multiGrid(first)) hipLaunchKernelGGL(multiTileFirstKernel, dim3(multiGrid(first)), dim3(kThreads), 0, s, first);
    if (multiGrid(tiles)) hipLaunchKernelGGL(multiTileBoundsKernel, dim3(multiGrid(tiles)), dim3(kThreads), 0, s, tiles);
    if (multiGrid(leaves))
        hipLaunchKernelGGL(multiTileCountsKernel, dim3(multiGrid(leaves)), dim3(kThreads), 0, s, leaves);
}

void multiPlan(hipStream_t s, const TreeDev *trees, int count, int phase, bool nf, std::uint32_t *start)
{
    hipLaunchKernelGGL(multiPlanKernel, dim3(1), dim3(kPlanThreads), 0, s, trees, count, phase, nf ? 1 : 0, start);
}
TreeSet plannedSet(const TreeDev *trees, int count, const std::uint32_t *start, PlanShape shape)
{
    TreeSet set;
    set.trees = trees;
    set.start = start + std::size_t(shape) * (std::size_t(count) + 1);
    set.count = count;
    set.total = set.start + count;
    return set;
}

std::size_t passTiles(std::size_t cap) { return cap / kPassTile + 1; }
std::size_t scanStatusWords(std::size_t items) { return items / kPassTile + 2; }

void octInitRoot(hipStream_t s, const TreeDev &t)
{
    hipLaunchKernelGGL(octInitKernel, dim3(clearBlocks(t.statusWords)), dim3(256), 0, s, t, t.statusWords);
}
void binInitRoot(hipStream_t s, const TreeDev &t)
{
    hipLaunchKernelGGL(binInitKernel, dim3(clearBlocks(t.statusWords)), dim3(256), 0, s, t, t.statusWords);
}
void updateReset(hipStream_t s, const TreeDev &t)
{
    hipLaunchKernelGGL(resetKernel, dim3(clearBlocks(t.statusWords)), dim3(256), 0, s, t.st, t.P, t.n, t.scanStatus,
                       t.statusWords, t.scanTicket, t.nfCount);
}

void octDecideScan(hipStream_t s, const TreeDev &t, const OctParams &p, long capacity, std::size_t gridLeaves)
{
    OctDecider d{{t.K[0], t.K[1]}, {t.N[0], t.N[1]}, {t.NF[0], t.NF[1]}, {t.gcm[0], t.gcm[1]}, p.bucket, p.nf};
    prof::Scope scope("decide_scan", s);
    hipLaunchKernelGGL((decideScanKernel<OctDecider>), dim3(unsigned(decideTiles(gridLeaves))), dim3(kDecideThreads),
                       0, s,
                       d, t.st, capacity, t.offs, t.scanStatus, t.scanTicket);
}
void octProcessCommit(hipStream_t s, const TreeDev &t, const OctParams &p, std::size_t gridLeaves)
{
    hipLaunchKernelGGL(octProcessKernel, dim3(ceilDiv(gridLeaves, kThreads)), dim3(kThreads), 0, s, t, p.nf);
    hipLaunchKernelGGL(commitKernel, dim3(clearBlocks(t.statusWords)), dim3(256), 0, s, t.st, 0, t.scanStatus,
                       t.statusWords, t.scanTicket, t.nfCount);
}
void octRefresh(hipStream_t s, const TreeDev &t, const OctParams &p, bool full, std::size_t gridLeaves)
{
    if (!p.nf)
    {
        prof::Scope scope("leaf_counts", s);
        hipLaunchKernelGGL(leafCountsKernel, dim3(ceilDiv(gridLeaves, kThreads)), dim3(kThreads), 0, s, t,
                           full ? 1 : 0);
        return;
    }
    prof::Scope scope("near_field", s);
    hipLaunchKernelGGL(nfCollectKernel, dim3(ceilDiv(gridLeaves, kThreads * kCollectRows)), dim3(kThreads), 0, s, t,
                       full ? 1 : 0);
    hipLaunchKernelGGL(nfComputeKernel, dim3(2048), dim3(kThreads), 0, s, t, p.bucket, p.hilbert ? 1 : 0);
    hipLaunchKernelGGL(nfGroupKernel, dim3(2048), dim3(kThreads), 0, s, t, p.hilbert ? 1 : 0, p.bucket);
}
void octPass(hipStream_t s, const TreeDev &t, const OctParams &p, long capacity, std::size_t gridLeaves)
{
    octDecideScan(s, t, p, capacity, gridLeaves);
    octProcessCommit(s, t, p, gridLeaves);
    octRefresh(s, t, p, false, gridLeaves);
}

void binDecideScan(hipStream_t s, const TreeDev &t, const BinParams &p, long capacity, std::size_t gridLeaves)
{
    BinDecider d{{t.K[0], t.K[1]}, {t.N[0], t.N[1]}, p.bucket, p.maxDepth};
    prof::Scope scope("decide_scan", s);
    hipLaunchKernelGGL((decideScanKernel<BinDecider>), dim3(unsigned(decideTiles(gridLeaves))), dim3(kDecideThreads),
                       0, s,
                       d, t.st, capacity, t.offs, t.scanStatus, t.scanTicket);
}
void binProcessCommit(hipStream_t s, const TreeDev &t, std::size_t gridLeaves)
{
    hipLaunchKernelGGL(binProcessKernel, dim3(ceilDiv(gridLeaves, kThreads)), dim3(kThreads), 0, s, t);
    hipLaunchKernelGGL(commitKernel, dim3(clearBlocks(t.statusWords)), dim3(256), 0, s, t.st, 1, t.scanStatus,
                       t.statusWords, t.scanTicket, t.nfCount);
}
// Tile b of P is [b T, (b+1) T). Leaf boundary i belongs to the tile that
// holds its lower bound in P: tile b owns the boundaries with
// P[bT - 1] < K[i] <= P[(b+1)T - 1] (the last tile also those above every
// key), i.e. tileFirst[b] <= i < tileFirst[b + 1].
void leafCountsByTiles(hipStream_t s, const TreeDev &t)
{
    prof::Scope scope("leaf_counts", s);
    const std::size_t tiles = (t.n + kBoundTile - 1) / kBoundTile;
    if (tiles)
    {
        hipLaunchKernelGGL(tileFirstKernel, dim3(ceilDiv(tiles + 1, kThreads)), dim3(kThreads), 0, s, t);
        hipLaunchKernelGGL(tileBoundsKernel, dim3(unsigned(tiles)), dim3(kThreads), 0, s, t);
    }
    hipLaunchKernelGGL(tileCountsKernel, dim3(2048), dim3(kThreads), 0, s, t);
}

__global__ __launch_bounds__(kThreads) void leafViolationsKernel(const Key *K, std::size_t L, const Key *P,
                                                                 std::size_t n, unsigned bucket, int bits,
                                                                 unsigned long long *counts)
{
    const std::size_t group = std::size_t(1) << bits;
    unsigned long long over = 0, merge = 0;
    for (std::size_t i = std::size_t(blockIdx.x) * kThreads + threadIdx.x; i < L;
         i += std::size_t(gridDim.x) * kThreads)
    {
        const Key range = K[i + 1] - K[i];
        const std::size_t a = lowerBound(P, n, K[i]);
        if (lowerBound(P, n, K[i + 1]) - a > bucket && range > 1) ++over;
        // 2^bits leaves spanning one aligned parent are exactly its children
        if (range < kMaxKey && i + group <= L)
        {
            const Key parent = range << bits;
            if (K[i] % parent == 0 && K[i + group] - K[i] == parent && lowerBound(P, n, K[i + group]) - a <= bucket)
                ++merge;
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
    {
        over += __shfl_xor(over, o, 64);
        merge += __shfl_xor(merge, o, 64);
    }
    if (__lane_id() == 0)
    {
        if (over) atomicAdd(&counts[0], over);
        if (merge) atomicAdd(&counts[1], merge);
    }
}

void leafCountViolations(hipStream_t s, const Key *K, std::size_t leaves, const Key *P, std::size_t n,
                         unsigned bucket, int bits, unsigned long long *counts)
{
    if (!leaves) return;
    const std::size_t want = (leaves + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(leafViolationsKernel, dim3(unsigned(want < 4096 ? want : 4096)), dim3(kThreads), 0, s, K,
                       leaves, P, n, bucket, bits, counts);
}

void leafCounts(hipStream_t s, const TreeDev &t, std::size_t gridLeaves)
{
    prof::Scope scope("leaf_counts", s);
    hipLaunchKernelGGL(leafCountsKernel, dim3(ceilDiv(gridLeaves, kThreads)), dim3(kThreads), 0, s, t, 1);
}
void binPass(hipStream_t s, const TreeDev &t, const BinParams &p, long capacity, std::size_t gridLeaves)
{
    binDecideScan(s, t, p, capacity, gridLeaves);
    binProcessCommit(s, t, gridLeaves); // the process step writes the new leaf counts
}

void binDirectCount(hipStream_t s, const TreeDev &t, const BinParams &p, BinDirectScratch &d)
{
    hipLaunchKernelGGL(resetKernel, dim3(1), dim3(64), 0, s, t.st, t.P, t.n, d.status, std::size_t(0), d.ticket,
                       nullptr);
    prof::Scope scope("bin_direct", s);
    const unsigned tiles = unsigned(t.n / kQ + 1);
    const int B = int(p.bucket);
    const std::size_t rowD = (std::size_t(kQ + B + 1) + 7) & ~std::size_t(7);
    const std::size_t lds = std::size_t(kQ + 2 * B + 2) * sizeof(Key) + 2 * rowD + kQ + 8;
    hipLaunchKernelGGL(binCountKernel, dim3(tiles), dim3(kThreads), lds, s, t.P, t.st, p.bucket, p.maxDepth, d.delta,
                       d.cnt, d.tileSum);
    hipLaunchKernelGGL(binScanTilesKernel, dim3(1), dim3(1024), 0, s, t.st, d.tileSum, d.total);
}
void binDirectWrite(hipStream_t s, const TreeDev &t, const BinParams &p, BinDirectScratch &d, long capacity)
{
    prof::Scope scope("bin_direct", s);
    hipLaunchKernelGGL(binEmitKernel, dim3(unsigned(t.n / kQ + 1)), dim3(kThreads), 0, s, t.P, t.st, d.delta, d.cnt,
                       d.tileSum, t.K[0], d.bound, d.total, capacity);
    hipLaunchKernelGGL(binBoundsToCountsKernel, dim3(ceilDiv(std::size_t(capacity), kThreads)), dim3(kThreads), 0, s,
                       d.bound, t.N[0], d.total, capacity);
    hipLaunchKernelGGL(binDirectStateKernel, dim3(1), dim3(1), 0, s, t.st, d.total);
}
} // namespace at
