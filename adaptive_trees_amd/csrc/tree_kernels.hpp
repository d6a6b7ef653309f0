// tree_kernels.hpp -- launchers of the octree / binary rebalance kernels.
#pragma once

#include "common.hpp"

namespace at
{
// Device buffers of one tree; [0]/[1] are the A/B halves selected by
// DevState::activeBuffer.
struct TreeDev
{
    Key *K[2] = {nullptr, nullptr};              // cornerstone, capacity + 1
    std::uint32_t *N[2] = {nullptr, nullptr};    // leaf counts
    std::uint32_t *NF[2] = {nullptr, nullptr};   // near-field counts (octree NF)
    std::uint8_t *gcm[2] = {nullptr, nullptr};   // groupCanMerge (octree NF)
    std::uint8_t *nfDirty = nullptr;
    std::uint8_t *countDirty = nullptr;          // LeafCount octree: leaf count still to be searched
    std::uint8_t *groupDirty = nullptr;
    std::int64_t *offs = nullptr;                // scanned node ops, capacity + 1
    std::uint64_t *scanStatus = nullptr;         // chained-scan status words
    std::uint32_t *scanTicket = nullptr;
    std::uint32_t *nfTasks = nullptr;            // (leaf << 1 | kind) work items
    std::uint32_t *nfGroups = nullptr;           // first leaf of each complete sibling group to refresh
    std::uint32_t *nfCount = nullptr;            // [0] nfTasks, [1] nfGroups
    std::size_t statusWords = 0;
    DevState *st = nullptr;
    const Key *P = nullptr;
    const Key *Pidx = nullptr;                   // sampled search index of P (indexSize(n) keys)
    std::uint32_t *tileFirst = nullptr;          // full leaf counts by tiles of P: n / kBoundTile + 2 entries
    std::size_t n = 0;                           // particle slots
    long capacity = 0;                           // leaf capacity of a pass (multi-tree launches)
    std::uint32_t *passLog = nullptr;            // device-driven batch builds: new leaf count of each committed pass
};

// Trees driven by one launch. Virtual block b belongs to the tree k with
// start[k] <= b < start[k+1]; start is a prefix of per-tree block counts (one
// table per launch shape), either computed on the host and copied to the
// device (blocks = start[count], one launched block per virtual block) or
// planned on the device from the trees' states (total points at
// start[count]; a fixed grid loops over the virtual blocks).
struct TreeSet
{
    const TreeDev *trees = nullptr;       // device array [count]
    const std::uint32_t *start = nullptr; // device array [count + 1]
    int count = 0;
    unsigned blocks = 0;                  // start[count] (host-planned tables)
    const std::uint32_t *total = nullptr; // device start[count] (device-planned tables)
};

// Launch shapes planned on the device between the steps of a pass.
enum PlanShape { kPlanTiles = 0, kPlanProcess, kPlanCommit, kPlanLeaves, kPlanNf, kPlanShapes };
// Per-tree block counts from the trees' device states, written as the
// prefixes start[shape * (count + 1) + k]. phase 0 (before a decision):
// decision tiles of the trees not converged. phase 1 (after it): process
// blocks over the old leaves, one commit block, refresh blocks over the new
// leaves and near-field work groups (nf), for the trees whose pass changed
// the tree and fitted its capacity.
void multiPlan(hipStream_t s, const TreeDev *trees, int count, int phase, bool nf, std::uint32_t *start);
TreeSet plannedSet(const TreeDev *trees, int count, const std::uint32_t *start, PlanShape shape);

// Per-tree block counts of each multi-tree launch shape.
unsigned clearBlocksFor(std::size_t statusWords); // status-word clearing (init, reset, commit)
unsigned tileBlocksFor(std::size_t gridLeaves);   // decision + chained scan tiles
unsigned leafBlocksFor(std::size_t gridLeaves);   // one thread per leaf, 8 rows of 256 leaves per block
unsigned nfBlocksFor(std::size_t gridLeaves);     // near-field work groups

constexpr int kPassTile = 2048;                  // leaves per chained-scan tile
std::size_t passTiles(std::size_t cap);
// Words of chained-scan status needed for `items` scanned elements.
std::size_t scanStatusWords(std::size_t items);

struct OctParams
{
    unsigned bucket;
    bool nf;
    bool hilbert;
};
struct BinParams
{
    unsigned bucket;
    unsigned maxDepth;
};

// ---- octree ----
void octInitRoot(hipStream_t s, const TreeDev &t);
void updateReset(hipStream_t s, const TreeDev &t);
// One split/merge pass: decision + chained scan + capacity check, scatter
// into the inactive half, commit, then the count refresh.
void octPass(hipStream_t s, const TreeDev &t, const OctParams &p, long capacity, std::size_t gridLeaves);
void octDecideScan(hipStream_t s, const TreeDev &t, const OctParams &p, long capacity, std::size_t gridLeaves);
void octProcessCommit(hipStream_t s, const TreeDev &t, const OctParams &p, std::size_t gridLeaves);
// full: recount every leaf (P changed); otherwise only the leaves the last
// process step could not derive.
void octRefresh(hipStream_t s, const TreeDev &t, const OctParams &p, bool full, std::size_t gridLeaves);

// ---- binary ----
void binInitRoot(hipStream_t s, const TreeDev &t);
void binPass(hipStream_t s, const TreeDev &t, const BinParams &p, long capacity, std::size_t gridLeaves);
void binDecideScan(hipStream_t s, const TreeDev &t, const BinParams &p, long capacity, std::size_t gridLeaves);
void binProcessCommit(hipStream_t s, const TreeDev &t, std::size_t gridLeaves);
void leafCounts(hipStream_t s, const TreeDev &t, std::size_t gridLeaves);
// Full leaf counts read P once, tile by tile (for small leaves, where a
// search per leaf would touch nearly every line of P anyway); needs
// t.tileFirst and uses t.offs as scratch.
constexpr std::size_t kBoundTile = 2048;
constexpr unsigned kTileCountBucket = 64; // leaf size up to which the tiled count is used
void leafCountsByTiles(hipStream_t s, const TreeDev &t);
// One-pass construction of the converged binary tree into K[0]/N[0]:
// phase 1 returns the number of internal nodes (the host sizes buffers),
// phase 2 writes the sorted leaf boundaries and their particle counts.
struct BinDirectScratch
{
    std::int8_t *delta = nullptr;       // per sorted particle: depth of its deepest internal node
    std::uint8_t *cnt = nullptr;        // per interval q in [0, nact]: boundaries in (P[q-1], P[q]]
    std::uint32_t *tileSum = nullptr;   // per tile: boundary count, then its exclusive prefix
    std::uint32_t *bound = nullptr;     // per boundary: index of its first particle
    unsigned long long *total = nullptr;
    std::uint64_t *status = nullptr;
    std::uint32_t *ticket = nullptr;
};
constexpr unsigned kDirectMaxBucket = 1024;
void binDirectCount(hipStream_t s, const TreeDev &t, const BinParams &p, BinDirectScratch &d);
// Emits K and N when the leaf count (total + 1, on the device) fits
// `capacity`, and writes the tree's state (its leaf count in any case).
void binDirectWrite(hipStream_t s, const TreeDev &t, const BinParams &p, BinDirectScratch &d, long capacity);

// Leaf-count convergence check (benchmarks/drivers/cstone_check.cu:17-43)
// over leaf boundaries K[0..leaves] and sorted keys P[0..n), populations
// recounted from P: counts[0] = leaves with more than `bucket` particles
// and a key range above 1; counts[1] = aligned groups of 2^bits sibling
// leaves (8 octree, 2 binary) that exactly tile their parent and hold at
// most `bucket` particles together. Adds to counts (device, two u64).
void leafCountViolations(hipStream_t s, const Key *K, std::size_t leaves, const Key *P, std::size_t n,
                         unsigned bucket, int bits, unsigned long long *counts);

// Builds t.Pidx from t.P (after every sort).
void buildSearchIndex(hipStream_t s, const TreeDev &t);
unsigned indexBlocksFor(std::size_t n);

// ---- multi-tree launches (segmented over a TreeSet) ----
void multiBuildSearchIndex(hipStream_t s, const TreeSet &index);
void multiOctInit(hipStream_t s, const TreeSet &clear);
void multiReset(hipStream_t s, const TreeSet &clear);
void multiClearScan(hipStream_t s, const TreeSet &clear);
void multiOctDecideScan(hipStream_t s, const TreeSet &tiles, const OctParams &p);
void multiOctProcess(hipStream_t s, const TreeSet &leaves, const OctParams &p);
void multiCommit(hipStream_t s, const TreeSet &clear, bool binary);
void multiOctRefresh(hipStream_t s, const TreeSet &leaves, const TreeSet &nf, const OctParams &p, bool full);
// Full leaf counts of LeafCount trees by tiles of P (leafCountsByTiles per
// member; t.tileFirst and t.offs as scratch): `first` with
// tileFirstBlocksFor(n) blocks per tree, `tiles` with tileBlocksOfP(n),
// `leaves` leaf-shaped.
unsigned tileFirstBlocksFor(std::size_t n);
unsigned tileBlocksOfP(std::size_t n);
void multiLeafCountsByTiles(hipStream_t s, const TreeSet &first, const TreeSet &tiles, const TreeSet &leaves);
} // namespace at
