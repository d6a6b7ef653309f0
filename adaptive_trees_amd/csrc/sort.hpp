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

// Tile shape chosen by the autotuner (items per thread); 16 by default.
int sortItemsPerThread();
void setSortItemsPerThread(int items);

// Device scratch of one sort instance, sized for `capacity` keys.
struct SortWorkspace
{
    Key *keysAlt = nullptr;        // ping-pong key buffer [capacity]
    std::uint32_t *valsAlt = nullptr;
    std::uint32_t *hist = nullptr; // [kSortPasses][kRadix] digit counts
    std::uint32_t *lookback = nullptr; // [2][maxTiles][kRadix] chained-scan status
    std::uint32_t *tickets = nullptr;  // [kSortPasses + 1] tile tickets
    std::size_t capacity = 0;
    std::size_t maxTiles = 0;
    bool values = false;

    void allocate(std::size_t cap, bool withValues);
    void release();
    ~SortWorkspace() { release(); }
};

std::size_t sortTiles(std::size_t n);

// Key generation kinds fused with the digit histogram.
enum class KeyKind { Morton, Hilbert, Binary };

template <class T>
struct KeygenParams
{
    const T *x, *y, *z;
    GridBox<T> grid;   // octree keys
    BinaryGeom geom;   // binary keys
};

// Writes keys of n particles into ws.keysAlt (perm is implicit: slot index)
// and builds all digit histograms; then sorts into (keysOut, valsOut) with
// valsOut[p] = original slot. Stable: equal keys keep ascending slots.
template <class T>
void keygenAndSort(hipStream_t s, KeyKind kind, const KeygenParams<T> &p, std::size_t n,
                   Key *keysOut, std::uint32_t *valsOut, SortWorkspace &ws);

// Stable sort of existing keys (and optional values) in place in keys/vals.
// valsIn == nullptr: values become the input positions 0..n-1.
void sortPairs(hipStream_t s, Key *keys, std::uint32_t *vals, std::size_t n, SortWorkspace &ws,
               bool valuesAreIndices);
void sortKeys(hipStream_t s, Key *keys, std::size_t n, SortWorkspace &ws);
} // namespace at
