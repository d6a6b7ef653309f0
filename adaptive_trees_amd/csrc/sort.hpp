// sort.hpp -- stable LSD radix sort of 64-bit keys (+ 32-bit values) for gfx950.
#pragma once

#include "common.hpp"

namespace at
{
// Digit passes of 8 bits over the full 64-bit key.
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortPasses = 64 / kRadixBits;
constexpr int kSortThreads = 256;

// Tile shape (threads per block, keys per thread) and target number of
// ranges, chosen by the autotuner; defaults 512, 12 and 2048.
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

// Device scratch of one sort instance, sized for `capacity` keys.
struct SortWorkspace
{
    Key *keysAlt = nullptr;            // ping-pong key buffer [capacity]
    std::uint32_t *valsAlt = nullptr;  // ping-pong value buffer
    std::uint8_t *digits = nullptr;    // next-pass digit per key [capacity]
    std::uint32_t *rangeHist = nullptr;  // [kRadix][rows] digit counts per range (or range part)
    std::uint32_t *rangeOff = nullptr;   // [kRadix][ranges] exclusive offset of each range within its digit
    std::uint32_t *digitTotal = nullptr; // [kRadix]
    std::size_t capacity = 0;
    std::size_t maxRanges = 0;
    bool values = false;

    void allocate(std::size_t cap, bool withValues);
    void release();
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
void sortPairs(hipStream_t s, Key *keys, std::uint32_t *vals, std::size_t n, SortWorkspace &ws,
               bool valuesAreIndices);
void sortKeys(hipStream_t s, Key *keys, std::size_t n, SortWorkspace &ws);
} // namespace at
