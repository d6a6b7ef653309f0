// sort.hip -- stable LSD radix sort (one sweep per 8-bit digit) for gfx950.
//
// Replaces the reference's cub::DeviceRadixSort::SortPairs calls
// (src/octree/octree.cu:301-304, src/kdtree3d/kdtree3d.cu:168-170,
// src/common/tree_view.cu:664-689): ascending by all 64 key bits, equal keys
// keep their input order.
//
// Design (wave64, 256 threads = 4 waves, ITEMS keys per thread):
//  * one up-front kernel builds the digit histograms of all 8 passes (fused
//    with key generation for particle keys, so keys are read once less);
//  * each pass is one kernel: a tile takes a ticket (tile order = start
//    order, which gives the chained scan forward progress), ranks its keys
//    per wave with 8 ballots per key (match of equal digits, stable in lane
//    order), publishes its per-digit counts and looks back across earlier
//    tiles for the exclusive prefix (status word = 2-bit flag | 30-bit count),
//    then stages keys (and values) through LDS so global writes leave in runs
//    of equal digits.
#include "sort.hpp"

namespace at
{
namespace
{
constexpr int kWaves = kSortThreads / 64;
constexpr std::uint32_t kFlagAgg = 1u << 30;
constexpr std::uint32_t kFlagPre = 2u << 30;
constexpr std::uint32_t kValMask = (1u << 30) - 1;

int g_items = 16;

__device__ inline unsigned laneId() { return __lane_id(); }

// Exclusive scan of one value per thread over a 256-thread block.
__device__ inline std::uint32_t blockExclusiveScan(std::uint32_t v, std::uint32_t *waveTotals,
                                                   std::uint32_t &total)
{
    const unsigned lane = laneId();
    const int wave = threadIdx.x >> 6;
    std::uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const std::uint32_t t = __shfl_up(inc, o, 64);
        if (lane >= unsigned(o)) inc += t;
    }
    if (lane == 63) waveTotals[wave] = inc;
    __syncthreads();
    std::uint32_t before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w)
    {
        const std::uint32_t t = waveTotals[w];
        if (w < wave) before += t;
        total += t;
    }
    __syncthreads();
    return before + inc - v;
}

template <int ITEMS, bool kVals, bool kIota>
__global__ __launch_bounds__(kSortThreads) void onesweepKernel(
    const Key *__restrict__ kin, Key *__restrict__ kout, const std::uint32_t *__restrict__ vin,
    std::uint32_t *__restrict__ vout, std::size_t n, const std::uint32_t *__restrict__ hist,
    std::uint32_t *lookback, std::uint32_t *lookbackNext, std::uint32_t *ticket, int shift)
{
    constexpr int kTile = ITEMS * kSortThreads;
    __shared__ union
    {
        Key keys[kTile];
        std::uint32_t vals[kTile];
    } stage;
    __shared__ std::uint32_t waveCount[kWaves][kRadix];
    __shared__ std::uint32_t adjust[kRadix];
    __shared__ std::uint32_t scratch[2 * kWaves];
    __shared__ std::uint32_t tileShared;

    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const unsigned lane = laneId();
    if (tid == 0) tileShared = atomicAdd(ticket, 1u);
#pragma unroll
    for (int w = 0; w < kWaves; ++w) waveCount[w][tid] = 0;
    __syncthreads();
    const std::size_t tile = tileShared;
    const std::size_t base = tile * kTile;
    // clear this tile's status row of the next pass
    lookbackNext[tile * kRadix + tid] = 0;

    // ---- load (wave-striped: item j of lane l is element j*64 + l of the wave's chunk)
    Key key[ITEMS];
    std::uint32_t val[ITEMS];
    std::uint32_t digit[ITEMS];
    const std::size_t waveBase = base + std::size_t(wave) * (ITEMS * 64);
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
    {
        const std::size_t idx = waveBase + std::size_t(j) * 64 + lane;
        const bool ok = idx < n;
        key[j] = ok ? kin[idx] : kInvalidKey;
        if (kVals) val[j] = kIota ? std::uint32_t(idx) : (ok ? vin[idx] : 0u);
        digit[j] = ok ? unsigned(key[j] >> shift) & (kRadix - 1) : kRadix - 1;
    }

    // ---- rank within the wave: lanes with equal digits via 8 ballots
    const std::uint64_t ltMask = (std::uint64_t(1) << lane) - 1;
    std::uint32_t rank[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
    {
        std::uint64_t match = ~std::uint64_t(0);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b)
        {
            const bool bit = (digit[j] >> b) & 1u;
            const std::uint64_t ballot = __ballot(bit);
            match &= bit ? ballot : ~ballot;
        }
        const std::uint32_t lower = __popcll(match & ltMask);
        const std::uint32_t cnt = __popcll(match);
        const std::uint32_t before = waveCount[wave][digit[j]];
        rank[j] = before + lower;
        if (lower == 0) waveCount[wave][digit[j]] = before + cnt;
    }
    __syncthreads();

    // ---- per digit (thread tid owns digit tid): wave offsets and tile count
    const int d = tid;
    std::uint32_t tileCount = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w)
    {
        const std::uint32_t c = waveCount[w][d];
        waveCount[w][d] = tileCount;
        tileCount += c;
    }

    // ---- chained scan across tiles for this digit
    std::uint32_t *status = lookback + tile * kRadix + d;
    std::uint32_t exclusive = 0;
    if (tile == 0)
        stRel(status, kFlagPre | tileCount);
    else
    {
        stRel(status, kFlagAgg | tileCount);
        std::size_t j = tile - 1;
        for (;;)
        {
            const std::uint32_t v = ldAcq(lookback + j * kRadix + d);
            const std::uint32_t flag = v & ~kValMask;
            if (flag == 0)
            {
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            exclusive += v & kValMask;
            if (flag == kFlagPre) break;
            --j;
        }
        stRel(status, kFlagPre | (exclusive + tileCount));
    }

    // ---- global digit base (exclusive scan of this pass's histogram) and the
    //      tile-local digit start (exclusive scan of the tile counts)
    std::uint32_t total;
    const std::uint32_t digitBase = blockExclusiveScan(hist[d], scratch, total);
    const std::uint32_t tileStart = blockExclusiveScan(tileCount, scratch + kWaves, total);
    adjust[d] = digitBase + exclusive - tileStart;
    // waveCount now holds wave offsets; fold the tile start in
#pragma unroll
    for (int w = 0; w < kWaves; ++w) waveCount[w][d] += tileStart;
    __syncthreads();

    // ---- stage keys in tile order of the output
    std::uint32_t local[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
    {
        local[j] = waveCount[wave][digit[j]] + rank[j];
        stage.keys[local[j]] = key[j];
    }
    __syncthreads();
    const std::size_t remaining = n - base;
    const int valid = remaining < std::size_t(kTile) ? int(remaining) : kTile;
    std::uint32_t dst[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
    {
        const int i = r * kSortThreads + tid;
        if (i < valid)
        {
            const Key k = stage.keys[i];
            const unsigned dg = unsigned(k >> shift) & (kRadix - 1);
            dst[r] = adjust[dg] + std::uint32_t(i);
            kout[dst[r]] = k;
        }
    }
    if (kVals)
    {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) stage.vals[local[j]] = val[j];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
        {
            const int i = r * kSortThreads + tid;
            if (i < valid) vout[dst[r]] = stage.vals[i];
        }
    }
}

// Digit histograms of all passes from keys in memory; also clears the first
// pass's chained-scan status rows.
__global__ __launch_bounds__(256) void histogramKernel(const Key *__restrict__ keys, std::size_t n,
                                                       std::uint32_t *hist, std::uint32_t *status0,
                                                       std::size_t statusWords)
{
    __shared__ std::uint32_t h[kSortPasses][kRadix];
    for (int i = threadIdx.x; i < kSortPasses * kRadix; i += blockDim.x) (&h[0][0])[i] = 0;
    __syncthreads();
    const std::size_t stride = std::size_t(gridDim.x) * blockDim.x;
    for (std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        const Key k = keys[i];
#pragma unroll
        for (int p = 0; p < kSortPasses; ++p) atomicAdd(&h[p][(k >> (p * kRadixBits)) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kSortPasses * kRadix; i += blockDim.x)
    {
        const std::uint32_t c = (&h[0][0])[i];
        if (c) atomicAdd(&hist[i], c);
    }
    for (std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < statusWords; i += stride)
        status0[i] = 0;
}

template <class T, KeyKind K>
__global__ __launch_bounds__(256) void keygenKernel(KeygenParams<T> p, std::size_t n, Key *__restrict__ keys,
                                                    std::uint32_t *hist, std::uint32_t *status0,
                                                    std::size_t statusWords)
{
    __shared__ std::uint32_t h[kSortPasses][kRadix];
    for (int i = threadIdx.x; i < kSortPasses * kRadix; i += blockDim.x) (&h[0][0])[i] = 0;
    __syncthreads();
    const std::size_t stride = std::size_t(gridDim.x) * blockDim.x;
    for (std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        const T x = p.x[i], y = p.y[i], z = p.z[i];
        Key k;
        if (K == KeyKind::Binary)
            k = binaryKey<T>(x, y, z, p.geom);
        else
            k = isEscaped<T>(x, y, z) ? kInvalidKey : gridKey<T>(x, y, z, p.grid, K == KeyKind::Hilbert);
        keys[i] = k;
#pragma unroll
        for (int q = 0; q < kSortPasses; ++q) atomicAdd(&h[q][(k >> (q * kRadixBits)) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kSortPasses * kRadix; i += blockDim.x)
    {
        const std::uint32_t c = (&h[0][0])[i];
        if (c) atomicAdd(&hist[i], c);
    }
    for (std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < statusWords; i += stride)
        status0[i] = 0;
}

__global__ void clearKernel(std::uint32_t *a, std::size_t n)
{
    for (std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += std::size_t(gridDim.x) * blockDim.x)
        a[i] = 0;
}

unsigned histogramBlocks(std::size_t n)
{
    const unsigned want = ceilDiv(n, 256 * 8);
    return want < 1 ? 1 : (want > 2048 ? 2048 : want);
}

template <int ITEMS>
void runPasses(hipStream_t s, Key *kA, Key *kB, std::uint32_t *vA, std::uint32_t *vB, bool vals, bool iota,
               std::size_t n, SortWorkspace &ws)
{
    constexpr std::size_t tileKeys = std::size_t(ITEMS) * kSortThreads;
    const std::size_t tiles = (n + tileKeys - 1) / tileKeys;
    for (int pass = 0; pass < kSortPasses; ++pass)
    {
        const Key *src = (pass & 1) ? kB : kA;
        Key *dst = (pass & 1) ? kA : kB;
        const std::uint32_t *vsrc = (pass & 1) ? vB : vA;
        std::uint32_t *vdst = (pass & 1) ? vA : vB;
        std::uint32_t *lb = ws.lookback + std::size_t(pass & 1) * ws.maxTiles * kRadix;
        std::uint32_t *lbNext = ws.lookback + std::size_t((pass + 1) & 1) * ws.maxTiles * kRadix;
        const std::uint32_t *hist = ws.hist + pass * kRadix;
        std::uint32_t *ticket = ws.tickets + pass;
        const int shift = pass * kRadixBits;
        if (!vals)
            hipLaunchKernelGGL((onesweepKernel<ITEMS, false, false>), dim3(unsigned(tiles)), dim3(kSortThreads), 0,
                               s, src, dst, nullptr, nullptr, n, hist, lb, lbNext, ticket, shift);
        else if (iota && pass == 0)
            hipLaunchKernelGGL((onesweepKernel<ITEMS, true, true>), dim3(unsigned(tiles)), dim3(kSortThreads), 0, s,
                               src, dst, nullptr, vdst, n, hist, lb, lbNext, ticket, shift);
        else
            hipLaunchKernelGGL((onesweepKernel<ITEMS, true, false>), dim3(unsigned(tiles)), dim3(kSortThreads), 0,
                               s, src, dst, vsrc, vdst, n, hist, lb, lbNext, ticket, shift);
    }
    AT_HIP(hipGetLastError());
}

void dispatchPasses(hipStream_t s, Key *kA, Key *kB, std::uint32_t *vA, std::uint32_t *vB, bool vals, bool iota,
                    std::size_t n, SortWorkspace &ws)
{
    switch (g_items)
    {
    case 8: runPasses<8>(s, kA, kB, vA, vB, vals, iota, n, ws); break;
    case 12: runPasses<12>(s, kA, kB, vA, vB, vals, iota, n, ws); break;
    default: runPasses<16>(s, kA, kB, vA, vB, vals, iota, n, ws); break;
    }
}

void clearCounters(hipStream_t s, SortWorkspace &ws)
{
    hipLaunchKernelGGL(clearKernel, dim3(1), dim3(256), 0, s, ws.hist, std::size_t(kSortPasses * kRadix));
    hipLaunchKernelGGL(clearKernel, dim3(1), dim3(64), 0, s, ws.tickets, std::size_t(kSortPasses + 1));
}
} // namespace

int sortItemsPerThread() { return g_items; }
void setSortItemsPerThread(int items)
{
    if (items != 8 && items != 12 && items != 16) throw Error(1, "sort items per thread must be 8, 12 or 16");
    g_items = items;
}

std::size_t sortTiles(std::size_t n)
{
    const std::size_t tile = std::size_t(8) * kSortThreads; // smallest tile: upper bound on the tile count
    return (n + tile - 1) / tile + 1;
}

void SortWorkspace::allocate(std::size_t cap, bool withValues)
{
    if (cap <= capacity && (!withValues || values)) return;
    release();
    capacity = cap < 1 ? 1 : cap;
    maxTiles = sortTiles(capacity);
    values = withValues;
    AT_HIP(hipMalloc(&keysAlt, capacity * sizeof(Key)));
    if (withValues) AT_HIP(hipMalloc(&valsAlt, capacity * sizeof(std::uint32_t)));
    AT_HIP(hipMalloc(&hist, kSortPasses * kRadix * sizeof(std::uint32_t)));
    AT_HIP(hipMalloc(&lookback, 2 * maxTiles * kRadix * sizeof(std::uint32_t)));
    AT_HIP(hipMalloc(&tickets, (kSortPasses + 1) * sizeof(std::uint32_t)));
}

void SortWorkspace::release()
{
    if (keysAlt) (void)hipFree(keysAlt);
    if (valsAlt) (void)hipFree(valsAlt);
    if (hist) (void)hipFree(hist);
    if (lookback) (void)hipFree(lookback);
    if (tickets) (void)hipFree(tickets);
    keysAlt = nullptr; valsAlt = nullptr; hist = nullptr; lookback = nullptr; tickets = nullptr;
    capacity = 0; maxTiles = 0; values = false;
}

template <class T>
void keygenAndSort(hipStream_t s, KeyKind kind, const KeygenParams<T> &p, std::size_t n, Key *keysOut,
                   std::uint32_t *valsOut, SortWorkspace &ws)
{
    if (n == 0) return;
    if (n >= (std::size_t(1) << 30)) throw Error(3, "radix sort supports fewer than 2^30 keys per tree");
    ws.allocate(n, true);
    clearCounters(s, ws);
    const std::size_t statusWords = ws.maxTiles * kRadix;
    const unsigned blocks = histogramBlocks(n);
    switch (kind)
    {
    case KeyKind::Morton:
        hipLaunchKernelGGL((keygenKernel<T, KeyKind::Morton>), dim3(blocks), dim3(256), 0, s, p, n, keysOut,
                           ws.hist, ws.lookback, statusWords);
        break;
    case KeyKind::Hilbert:
        hipLaunchKernelGGL((keygenKernel<T, KeyKind::Hilbert>), dim3(blocks), dim3(256), 0, s, p, n, keysOut,
                           ws.hist, ws.lookback, statusWords);
        break;
    case KeyKind::Binary:
        hipLaunchKernelGGL((keygenKernel<T, KeyKind::Binary>), dim3(blocks), dim3(256), 0, s, p, n, keysOut,
                           ws.hist, ws.lookback, statusWords);
        break;
    }
    // an even number of ping-pong passes starting from keysOut ends in keysOut;
    // pass 0 generates the slot indices as values
    dispatchPasses(s, keysOut, ws.keysAlt, valsOut, ws.valsAlt, true, true, n, ws);
}

void sortPairs(hipStream_t s, Key *keys, std::uint32_t *vals, std::size_t n, SortWorkspace &ws, bool iota)
{
    if (n == 0) return;
    if (n >= (std::size_t(1) << 30)) throw Error(3, "radix sort supports fewer than 2^30 keys");
    ws.allocate(n, true);
    clearCounters(s, ws);
    hipLaunchKernelGGL(histogramKernel, dim3(histogramBlocks(n)), dim3(256), 0, s, keys, n, ws.hist, ws.lookback,
                       ws.maxTiles * kRadix);
    // ping-pong (keys, keysAlt): pass 0 keys -> alt, pass 7 alt -> keys
    dispatchPasses(s, keys, ws.keysAlt, vals, ws.valsAlt, true, iota, n, ws);
}

void sortKeys(hipStream_t s, Key *keys, std::size_t n, SortWorkspace &ws)
{
    if (n == 0) return;
    if (n >= (std::size_t(1) << 30)) throw Error(3, "radix sort supports fewer than 2^30 keys");
    ws.allocate(n, false);
    clearCounters(s, ws);
    hipLaunchKernelGGL(histogramKernel, dim3(histogramBlocks(n)), dim3(256), 0, s, keys, n, ws.hist, ws.lookback,
                       ws.maxTiles * kRadix);
    dispatchPasses(s, keys, ws.keysAlt, nullptr, nullptr, false, false, n, ws);
}

template void keygenAndSort<float>(hipStream_t, KeyKind, const KeygenParams<float> &, std::size_t, Key *,
                                   std::uint32_t *, SortWorkspace &);
template void keygenAndSort<double>(hipStream_t, KeyKind, const KeygenParams<double> &, std::size_t, Key *,
                                    std::uint32_t *, SortWorkspace &);
} // namespace at
