// sort.hpp -- stable radix sort of 64-bit keys (+ 32-bit values) for gfx950.
#pragma once

#include <string>
#include <vector>

#include "buffers.hpp"
#include "common.hpp"

namespace at
{
// Digit passes of 8 bits over the full 64-bit key.
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortPasses = 64 / kRadixBits;
constexpr int kMaxTileKeys = 8192; // scatter tile (threads x keys per thread) bound by LDS
constexpr int kSortThreads = 256;

// Tile shape (threads per block, keys per thread) and target number of
// ranges, chosen by the autotuner; defaults 512, 8 (no prefetch) and 8192.
int sortItemsPerThread();
void setSortItemsPerThread(int items);
int sortTargetRanges();
void setSortTargetRanges(int ranges);
int sortThreads();
void setSortThreads(int threads);
// 1: each scatter also writes the next pass's digit bytes and the next
// histogram reads those; 0: the histogram reads the keys.
int sortDigitArray();
void setSortDigitArray(int on);
// 1: the scatter loads the next tile during the current one (tile shapes
// 256/512 x 8/12; others ignore it).
int sortPrefetch();
void setSortPrefetch(int on);
// 1: sorts of at least sortMsdMinKeys() (key, value) pairs partition by the
// high digits first (MSD levels) and finish buckets of up to 2048 keys in
// LDS; 0: LSD passes over all 8 digits. Results are identical.
int sortMsd();
void setSortMsd(int on);
// The MSD threshold (default 4096 pairs; at least 2048).
std::size_t sortMsdMinKeys();
void setSortMsdMinKeys(std::size_t n);
// Digit width of the LSD passes of sorts below 2^20 keys that do not take
// the MSD levels: 8 (8 passes) or 11 (6 passes of 2048 digits). Results are
// identical.
int sortDigitBits();
void setSortDigitBits(int bits);

// Times every tile shape, then range count, then histogram source on a
// stable sort of n random 63-bit (key, index) pairs, then the LSD digit
// width on an LSD sort of min(n, 2^20 - 1) keys, and keeps the fastest
// settings (process-wide). Returns them as a JSON object.
std::string autotuneSort(std::size_t n, hipStream_t s);

// Partition of n keys into contiguous ranges of whole tiles; each scatter
// block owns one range and walks it tile by tile.
struct SortPlan
{
    std::size_t n = 0;
    int items = 16;
    int threads = 256;
    std::size_t tileKeys = 0;
    std::size_t rangeTiles = 1;
    std::size_t rangeKeys = 0;
    std::size_t ranges = 0;
    int parts = 1;          // histogram blocks per range
};
SortPlan makeSortPlan(std::size_t n);

// A contiguous range of keys within a segment (segmented sorts and MSD levels).
struct SegRange
{
    std::uint64_t begin;    // first key of the range in the concatenated array
    std::uint32_t len;      // keys in the range
    std::uint32_t localBase; // slot of the first key within its segment
    std::uint32_t seg;
    std::uint32_t pad;
};

// ---- MSD sort state (device) ----
// A level's segment: a bucket of more than kMsdChunk keys that the level
// partitions by its digit; its ranges are firstRange .. firstRange+ranges-1.
struct MsdSeg
{
    std::uint64_t begin;
    std::uint32_t size;
    std::uint32_t firstRange;
    std::uint32_t ranges;
    std::uint32_t pad;
};
// A run of consecutive finished buckets (at most kMsdChunk keys, or any
// size when bits == 0) sorted in LDS on its low `bits` bits; parity 1 = it
// lies in the odd levels' output buffer.
struct MsdChunk
{
    std::uint64_t begin;
    std::uint32_t len;
    std::uint8_t bits;
    std::uint8_t parity;
    std::uint16_t pad;
};
struct MsdCounters
{
    std::uint32_t segs[2];   // segments of the levels using list set 0 / 1
    std::uint32_t ranges[2];
    std::uint32_t chunks;
    std::uint32_t overflow;  // list capacity exceeded (never, by the bounds used)
    std::uint32_t pad[2];    // pad[0]: chunks taken by an early chunk-sort launch
};

// Device scratch of one sort instance, sized for `capacity` keys.
struct SortWorkspace
{
    Key *keysAlt = nullptr;            // ping-pong key buffer [capacity]
    std::uint32_t *valsAlt = nullptr;  // ping-pong value buffer
    std::uint8_t *digits = nullptr;    // next-pass digit per key [capacity]
    std::uint32_t *rangeHist = nullptr;  // [kRadix][rows] digit counts per range (or range part)
    std::uint32_t *rangeOff = nullptr;   // [kRadix][ranges] exclusive offset of each range within its digit
    std::uint32_t *digitTotal = nullptr; // [kRadix]
    // 11-bit LSD passes (workspaces of fewer than 2^20 keys): [2048][wideRanges]
    // histograms and offsets, digit totals [2048]
    std::uint32_t *wideHist = nullptr, *wideOff = nullptr, *wideTotal = nullptr;
    std::size_t wideRanges = 0;
    std::size_t capacity = 0;
    std::size_t maxRanges = 0;
    bool values = false;
    // MSD levels (allocated on first use)
    MsdCounters *msdCnt = nullptr;
    MsdSeg *msdRoot = nullptr;
    MsdSeg *msdSegs[2] = {nullptr, nullptr};
    SegRange *msdRanges[2] = {nullptr, nullptr};
    std::uint8_t *msdFirst[2] = {nullptr, nullptr}; // 1 for the first range of a segment
    std::uint32_t *msdHist = nullptr;      // [kRadix][msdStride]
    std::uint32_t *msdOff = nullptr;       // [kRadix][msdStride]
    std::uint32_t *msdSubTotals = nullptr; // [msdMaxSegs][kRadix]
    MsdChunk *msdChunks = nullptr;
    std::uint32_t *msdHostCount = nullptr; // pinned: a level's segment count read by the host
    std::size_t msdCapacity = 0, msdMaxSegs = 0, msdStride = 0, msdMaxChunks = 0, msdParents = 0;
    bool borrowed = false; // pointers into memory the workspace does not own (release() only forgets them)

    void allocate(std::size_t cap, bool withValues, hipStream_t s);
    // MSD lists for sorts of up to cap keys whose level 1 has `parents`
    // parent buckets besides the whole array (segmented sorts)
    void allocateMsd(std::size_t cap, std::size_t parents, hipStream_t s);
    // release(): the owner has synchronised its stream (destruction);
    // releaseOn(s): frees in stream s's order (regrowth while s may still
    // use the buffers).
    void release();
    void releaseOn(hipStream_t s);
    ~SortWorkspace() { release(); }
};

// Key generation kinds fused with the first pass's range histograms.
enum class KeyKind { Morton, Hilbert, Binary };

template <class T>
struct KeygenParams
{
    const T *x, *y, *z;
    GridBox<T> grid;   // octree keys
    BinaryGeom geom;   // binary keys
};

// Writes keys of n particles into keysOut (values are the slot indices) and
// sorts (keysOut, valsOut) in place with valsOut[p] = original slot. Stable:
// equal keys keep ascending slots.
template <class T>
void keygenAndSort(hipStream_t s, KeyKind kind, const KeygenParams<T> &p, std::size_t n,
                   Key *keysOut, std::uint32_t *valsOut, SortWorkspace &ws);

// Stable sort of existing keys (and values) in place. valuesAreIndices:
// values become the input positions 0..n-1 (vals is output only).
// msdAllowed = false keeps the LSD passes at any size (sorts of a few
// million keys with long runs of equal high digits, where the MSD levels'
// fixed launches cost more than they save).
void sortPairs(hipStream_t s, Key *keys, std::uint32_t *vals, std::size_t n, SortWorkspace &ws,
               bool valuesAreIndices, bool msdAllowed = true);
void sortKeys(hipStream_t s, Key *keys, std::size_t n, SortWorkspace &ws);

// Stable sort of at most kNoAllocMaxKeys pairs, like sortPairs, whose
// workspace is a fixed device array of the library (one per GPU, held until
// the process ends): no allocation of any kind. Asynchronous on `s`; every
// call on one GPU shares the array, so calls that overlap on one GPU race.
constexpr std::size_t kNoAllocMaxKeys = std::size_t(1) << 20;
void sortPairsNoAlloc(hipStream_t s, Key *keys, std::uint32_t *vals, std::size_t n, bool valuesAreIndices);
std::size_t noAllocScratchBytes();

// ---- segmented sort: many independent key arrays in one set of launches ----
// Segment s holds counts[s] particles at [segBegin[s], segBegin[s+1]) of one
// concatenated key array and is sorted on its own (stable, values = the
// particle's slot within its segment). Every stage is one launch over all
// segments: ranges never straddle a segment, the per-segment offset scan
// runs one block per segment.

template <class T>
struct SegCoords
{
    const T *x, *y, *z;
};

class SegmentedSort
{
public:
    // Plans ranges for the given segment sizes and allocates scratch.
    void plan(const std::vector<std::size_t> &counts, hipStream_t s);
    std::size_t total() const { return total_; }
    std::size_t segBegin(std::size_t seg) const { return segBegin_[seg]; }
    std::size_t ranges() const { return ranges_; }
    // Keys of every segment (particle j of segment s from coords[s]) into
    // keys[segBegin[s] + ...], sorted per segment with vals = slot in segment.
    // hostPoll = false: the MSD levels keep their fixed launch sequence (no
    // host read of a level count).
    template <class T>
    void keygenAndSort(hipStream_t s, KeyKind kind, const KeygenParams<T> &shared, const SegCoords<T> *coords,
                       Key *keys, std::uint32_t *vals, bool hostPoll = true);

private:
    std::size_t total_ = 0, ranges_ = 0, segments_ = 0, rangeKeys_ = 0;
    std::vector<std::size_t> segBegin_;
    DevBuf<SegRange> range_;
    DevBuf<std::uint32_t> segRange_;   // first range of each segment [segments + 1]
    DevBuf<std::uint64_t> segBeginD_;  // [segments + 1]
    DevBuf<std::uint32_t> hist_, off_; // [ranges][kRadix]
    DevBuf<Key> keysAlt_;
    DevBuf<std::uint32_t> valsAlt_;
    DevBuf<std::uint8_t> digits_;
    // MSD path (totals of at least 2^20 keys): the segments are level 1's parents
    bool msdReady_ = false;
    SortWorkspace msd_;
    DevBuf<MsdSeg> parents_;              // [segments]
    DevBuf<std::uint32_t> parentCount_;   // [1]
    DevBuf<std::uint32_t> segTotals_;     // [segments][kRadix]
};
} // namespace at
