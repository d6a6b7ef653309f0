// sort.hip -- stable radix sort for gfx950: MSD levels + LDS chunk sort for
// large sorts, LSD passes (8-bit digits, 8 passes) otherwise.
//
// Replaces the reference's cub::DeviceRadixSort::SortPairs calls
// (src/octree/octree.cu:301-304, src/kdtree3d/kdtree3d.cu:168-170,
// src/common/tree_view.cu:664-689): ascending by all 64 key bits, equal keys
// keep their input order.
//
// An LSD pass is range-partitioned, with no communication between blocks:
//   upsweep   digit histogram of each contiguous range of keys. Pass 0's
//             histograms come out of key generation itself; pass k+1 reads a
//             one-byte digit array that pass k's scatter wrote, not the keys.
//   scan      one block per digit: the digit's exclusive prefix over the
//             ranges and its total. Histograms are stored digit-major
//             (hist[digit][row]), so each block reads one contiguous row set;
//             every scatter block forms the digit bases from the 256 totals.
//   scatter   one block per range walks its range tile by tile (wave64,
//             ITEMS keys per thread). Keys are ranked per wave with 8 ballots
//             each (lanes with equal digits, stable in lane order), staged in
//             LDS in output order, and written in runs of equal digits at a
//             per-digit running offset kept in LDS.
// The MSD sort (further below) uses the same scatter for its partition
// levels and finishes small buckets in LDS.
// A chained (look-back) scan is avoided on purpose: with ~768 tiles resident
// on the 256 CUs the prefix chain made the look-back cost ~45% of a pass.
#include "sort.hpp"

#include <chrono>
#include <mutex>
#include <string>
#include <type_traits>

#include "profile.hpp"

namespace at
{
namespace
{
constexpr int kWaves = kSortThreads / 64;
constexpr std::size_t kHistogramBlocks = 4096;
constexpr std::size_t kXcds = 8; // accelerator complex dies of one MI355X, each with its own L2

int g_items = 8;
int g_threads = 512;
int g_targetRanges = 8192; // one 4096-key tile per range at 2^25 keys
int g_digitArray = 1; // scatter writes the next pass's digit bytes for the histogram
int g_prefetch = 0;   // scatter loads the next tile while working on the current one
int g_msd = 1;        // MSD levels + LDS chunk sort for large (key, value) sorts
int g_digitBits = 8;  // LSD digit width below the MSD threshold (8 or 11)
// sorts of at least this many (key, value) pairs take the MSD levels: from
// 4096 keys a partition level and the LDS chunk sort (about 14 launches and
// one host read) beat eight LSD passes of three launches each (requests of
// 1e4..1e6 particles 25-30% faster on a mixed serving trace)
std::size_t g_msdMinKeys = 4096;
// workspaces below this capacity hold the 11-bit LSD histograms
constexpr std::size_t kWideMaxKeys = std::size_t(1) << 20;

__device__ inline unsigned laneId() { return __lane_id(); }

// Workgroups are dispatched round-robin over the 8 XCDs, so block b runs on
// XCD b % 8. This maps blocks to rows (key ranges) so that each XCD gets a
// contiguous run of rows, in order: neighbouring rows write adjacent words
// of the digit-major histograms and adjacent pieces of every digit's output,
// and running them at the same time on one L2 lets the partial cache lines
// at those seams merge there before they are written back.
__device__ inline std::size_t xcdRow(std::size_t b, std::size_t rows)
{
    const std::size_t perXcd = rows / kXcds;
    return b < kXcds * perXcd ? (b % kXcds) * perXcd + b / kXcds : b;
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ inline std::uint32_t blockExclusiveScan(std::uint32_t v, std::uint32_t *waveTotals)
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
#pragma unroll
    for (int w = 0; w < kWaves; ++w)
        if (w < wave) before += waveTotals[w];
    __syncthreads();
    return before + inc - v;
}

// Adds a digit to a block histogram. A digit equal on every active lane of
// the wave (common for the high digits of clustered data) is added once
// instead of by 64 conflicting LDS atomics.
__device__ inline void countDigit(std::uint32_t *h, unsigned d)
{
    const std::uint64_t active = __ballot(1);
    const int leader = __ffsll((long long)active) - 1;
    const unsigned d0 = __builtin_amdgcn_readfirstlane(d); // the leader's digit, read by the scalar unit
    if (__ballot(d != d0) == 0)
    {
        if (int(laneId()) == leader) atomicAdd(&h[d0], std::uint32_t(__popcll(active)));
    }
    else
        atomicAdd(&h[d], 1u);
}

// Block histogram -> row blockIdx.x of the digit-major histogram table
// (hist[digit * gridDim.x + row]).
__device__ inline void flushHistogram(const std::uint32_t *h, std::uint32_t *hist)
{
    __syncthreads();
    hist[std::size_t(threadIdx.x) * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

constexpr int kKeygenUnroll = 4;

template <class T, KeyKind K>
__device__ inline Key coordKey(const KeygenParams<T> &p, T x, T y, T z)
{
    if (K == KeyKind::Binary) return binaryKey<T>(x, y, z, p.geom);
    return isEscaped<T>(x, y, z) ? kInvalidKey : gridKey<T>(x, y, z, p.grid, K == KeyKind::Hilbert);
}

template <class T, KeyKind K>
__device__ inline Key particleKey(const KeygenParams<T> &p, std::size_t i)
{
    return coordKey<T, K>(p, p.x[i], p.y[i], p.z[i]);
}

// Four consecutive coordinates from a 16-byte aligned address.
__device__ inline void load4(const float *a, std::size_t i, float v[4])
{
    const float4 q = *reinterpret_cast<const float4 *>(a + i);
    v[0] = q.x, v[1] = q.y, v[2] = q.z, v[3] = q.w;
}
__device__ inline void load4(const double *a, std::size_t i, double v[4])
{
    const double2 q0 = reinterpret_cast<const double2 *>(a + i)[0];
    const double2 q1 = reinterpret_cast<const double2 *>(a + i)[1];
    v[0] = q0.x, v[1] = q0.y, v[2] = q1.x, v[3] = q1.y;
}

// The Hilbert encode table (hilbert_table::kTables.enc) staged in LDS by a
// block that makes Hilbert keys; all threads call it.
constexpr int kHilbertRows = 48 * 8;
__device__ inline void stageHilbertEnc(std::uint16_t *lds)
{
    for (int i = threadIdx.x; i < kHilbertRows; i += blockDim.x) lds[i] = hilbert_table::kTables.enc[i];
}

template <class T, KeyKind K, bool FAST>
__device__ inline Key keyOf(const KeygenParams<T> &p, const std::uint32_t *spread, const std::uint16_t *henc, T x,
                            T y, T z)
{
    if constexpr (FAST) return binaryKeyFast<T>(x, y, z, p.geom, spread);
    else if constexpr (K == KeyKind::Hilbert)
        return isEscaped<T>(x, y, z) ? kInvalidKey : gridKeyHilbertTable<T>(x, y, z, p.grid, henc);
    else return coordKey<T, K>(p, x, y, z);
}

// Keys of particles [begin, end) into keys, their digits at histShift
// counted in h. With 16-byte aligned coordinate arrays each thread takes
// groups of 4 consecutive particles (16-byte loads and stores), several
// groups in flight; otherwise kKeygenUnroll single particles.
template <class T, KeyKind K, bool FAST>
__device__ __forceinline__ void keygenPart(const KeygenParams<T> &p, const std::uint32_t *spread,
                                           const std::uint16_t *henc, std::uint32_t *h, Key *__restrict__ keys,
                                           std::size_t begin, std::size_t end, int histShift)
{
    const bool vec = ((reinterpret_cast<std::uintptr_t>(p.x) | reinterpret_cast<std::uintptr_t>(p.y) |
                       reinterpret_cast<std::uintptr_t>(p.z) | reinterpret_cast<std::uintptr_t>(keys)) & 15) == 0 &&
                     begin % 4 == 0;
    std::size_t i;
    if (vec)
    {
        // FP32: 16 particles (12 float4 loads in flight) per thread and step,
        // so a 4096-particle range is one step
        constexpr int kG = sizeof(T) == 4 ? 4 : 2;
        constexpr std::size_t kGroupStep = std::size_t(kSortThreads) * 4;
        // block-uniform steps of kG * kGroupStep particles, then the tail
        std::size_t base = begin;
        for (; base + kG * kGroupStep <= end; base += kG * kGroupStep)
        {
            i = base + 4 * std::size_t(threadIdx.x);
            T x[kG][4], y[kG][4], z[kG][4];
#pragma unroll
            for (int g = 0; g < kG; ++g)
            {
                load4(p.x, i + g * kGroupStep, x[g]);
                load4(p.y, i + g * kGroupStep, y[g]);
                load4(p.z, i + g * kGroupStep, z[g]);
            }
#pragma unroll
            for (int g = 0; g < kG; ++g)
            {
                Key k[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) k[u] = keyOf<T, K, FAST>(p, spread, henc, x[g][u], y[g][u], z[g][u]);
                ulonglong2 *dst = reinterpret_cast<ulonglong2 *>(keys + i + g * kGroupStep);
                dst[0] = make_ulonglong2(k[0], k[1]);
                dst[1] = make_ulonglong2(k[2], k[3]);
#pragma unroll
                for (int u = 0; u < 4; ++u) atomicAdd(&h[unsigned(k[u] >> histShift) & (kRadix - 1)], 1u);
            }
        }
        i = base + threadIdx.x; // the rest of the part, one particle per thread
    }
    else
    {
        constexpr std::size_t kStep = std::size_t(kSortThreads) * kKeygenUnroll;
        i = begin + threadIdx.x;
        for (; i + (kKeygenUnroll - 1) * kSortThreads < end; i += kStep)
        {
            Key k[kKeygenUnroll];
#pragma unroll
            for (int u = 0; u < kKeygenUnroll; ++u)
            {
                const std::size_t j = i + std::size_t(u) * kSortThreads;
                k[u] = keyOf<T, K, FAST>(p, spread, henc, p.x[j], p.y[j], p.z[j]);
            }
#pragma unroll
            for (int u = 0; u < kKeygenUnroll; ++u)
            {
                keys[i + std::size_t(u) * kSortThreads] = k[u];
                atomicAdd(&h[unsigned(k[u] >> histShift) & (kRadix - 1)], 1u);
            }
        }
    }
    for (; i < end; i += kSortThreads)
    {
        const Key k = keyOf<T, K, FAST>(p, spread, henc, p.x[i], p.y[i], p.z[i]);
        keys[i] = k;
        atomicAdd(&h[unsigned(k >> histShift) & (kRadix - 1)], 1u);
    }
}

// Part `blockIdx.x % parts` of a range: its keys and a partial digit-0
// histogram (parts start at multiples of 4). Binary keys take the
// branch-free form when the geometry allows it (a block-uniform choice).
template <class T, KeyKind K>
__global__ __launch_bounds__(kSortThreads) void keygenKernel(KeygenParams<T> p, std::size_t n, std::size_t rangeKeys,
                                                             int parts, Key *__restrict__ keys, std::uint32_t *partHist,
                                                             int histShift)
{
    static_assert(kSortThreads == kRadix, "one table entry and one digit per thread");
    // one histogram per wave: the lowest digit is close to uniform, so plain
    // LDS atomics on a wave-private table collide least
    __shared__ std::uint32_t hw[kWaves][kRadix];
    __shared__ std::uint32_t spread[K == KeyKind::Binary ? kRadix : 1]; // spread3 of every byte
    __shared__ std::uint16_t henc[K == KeyKind::Hilbert ? kHilbertRows : 1];
#pragma unroll
    for (int w = 0; w < kWaves; ++w) hw[w][threadIdx.x] = 0;
    if constexpr (K == KeyKind::Binary) spread[threadIdx.x] = std::uint32_t(spread3(threadIdx.x));
    if constexpr (K == KeyKind::Hilbert) stageHilbertEnc(henc);
    __syncthreads();
    std::uint32_t *h = hw[threadIdx.x >> 6];
    const std::size_t row = xcdRow(blockIdx.x, gridDim.x);
    const std::size_t r = row / unsigned(parts), part = row % unsigned(parts);
    const std::size_t partKeys = rangeKeys / std::size_t(parts);
    const std::size_t rEnd = (r + 1) * rangeKeys < n ? (r + 1) * rangeKeys : n;
    const std::size_t begin = r * rangeKeys + part * partKeys;
    const std::size_t end0 = begin + partKeys;
    const std::size_t end = end0 < rEnd ? end0 : rEnd;
    if constexpr (K == KeyKind::Binary)
    {
        if (p.geom.fast) keygenPart<T, K, true>(p, spread, henc, h, keys, begin, end, histShift);
        else keygenPart<T, K, false>(p, spread, henc, h, keys, begin, end, histShift);
    }
    else
        keygenPart<T, K, false>(p, spread, henc, h, keys, begin, end, histShift);
    __syncthreads();
    std::uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) t += hw[w][threadIdx.x];
    partHist[std::size_t(threadIdx.x) * gridDim.x + row] = t;
}

// Digit histogram of range blockIdx.x read from the keys: two keys per
// 16-byte load when the array is 16-byte aligned (ranges start at even
// indices), one LDS histogram per wave.
__global__ __launch_bounds__(kSortThreads) void upsweepKeysKernel(const Key *__restrict__ keys, std::size_t n,
                                                                  std::size_t rangeKeys, int shift,
                                                                  std::uint32_t *rangeHist)
{
    __shared__ std::uint32_t h[kWaves][kRadix];
#pragma unroll
    for (int w = 0; w < kWaves; ++w) h[w][threadIdx.x] = 0;
    __syncthreads();
    std::uint32_t *mine = h[threadIdx.x >> 6];
    const std::size_t row = xcdRow(blockIdx.x, gridDim.x);
    const std::size_t begin = row * rangeKeys;
    const std::size_t end = begin + rangeKeys < n ? begin + rangeKeys : n;
    if ((reinterpret_cast<std::uintptr_t>(keys) & 15) == 0)
    {
        const ulonglong2 *pairs = reinterpret_cast<const ulonglong2 *>(keys);
        constexpr int kU = 2;
        std::size_t i = begin + 2 * threadIdx.x;
        for (; i + 2 * kSortThreads * (kU - 1) + 1 < end; i += 2 * kSortThreads * kU)
        {
            ulonglong2 q[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) q[u] = pairs[(i >> 1) + std::size_t(u) * kSortThreads];
#pragma unroll
            for (int u = 0; u < kU; ++u)
            {
                countDigit(mine, unsigned(q[u].x >> shift) & (kRadix - 1));
                countDigit(mine, unsigned(q[u].y >> shift) & (kRadix - 1));
            }
        }
        for (; i < end; i += 2 * kSortThreads)
        {
            atomicAdd(&mine[unsigned(keys[i] >> shift) & (kRadix - 1)], 1u);
            if (i + 1 < end) atomicAdd(&mine[unsigned(keys[i + 1] >> shift) & (kRadix - 1)], 1u);
        }
    }
    else
        for (std::size_t i = begin + threadIdx.x; i < end; i += kSortThreads)
            atomicAdd(&mine[unsigned(keys[i] >> shift) & (kRadix - 1)], 1u);
    __syncthreads();
    std::uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) t += h[w][threadIdx.x];
    rangeHist[std::size_t(threadIdx.x) * gridDim.x + row] = t;
}

// Digit histograms of kUpRows consecutive ranges from the one-byte digit
// array: one wave per range, 16 digits per load (ranges start at multiples
// of the tile size), one LDS histogram per wave. Thread d then writes digit
// d's counts of the block's ranges, adjacent in the digit-major table, as one
// 32-byte piece.
constexpr int kUpRows = 8;
__global__ __launch_bounds__(kUpRows * 64) void upsweepDigitsKernel(const std::uint8_t *__restrict__ digits,
                                                                    std::size_t n, std::size_t rangeKeys,
                                                                    std::size_t ranges, std::uint32_t *rangeHist)
{
    __shared__ std::uint32_t h[kUpRows][kRadix];
    for (int i = threadIdx.x; i < kUpRows * kRadix; i += kUpRows * 64) (&h[0][0])[i] = 0;
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    const unsigned lane = laneId();
    const std::size_t r0 = xcdRow(blockIdx.x, gridDim.x) * kUpRows;
    const std::size_t r = r0 + std::size_t(wave);
    std::uint32_t *mine = h[wave];
    if (r < ranges)
    {
        const std::size_t begin = r * rangeKeys;
        const std::size_t end = begin + rangeKeys < n ? begin + rangeKeys : n;
        const uint4 *quads = reinterpret_cast<const uint4 *>(digits);
        for (std::size_t i = begin + 16 * lane; i < end; i += 16 * 64)
        {
            if (i + 16 <= end)
            {
                const uint4 q = quads[i >> 4];
                const std::uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
                for (int k = 0; k < 4; ++k)
#pragma unroll
                    for (int b = 0; b < 4; ++b) countDigit(mine, (w[k] >> (8 * b)) & 0xffu);
            }
            else
                for (std::size_t j = i; j < end; ++j) atomicAdd(&mine[digits[j]], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < kRadix)
    {
        const unsigned d = threadIdx.x;
        std::uint32_t *out = rangeHist + std::size_t(d) * ranges + r0;
        if (r0 + kUpRows <= ranges && ranges % 4 == 0)
        {
            reinterpret_cast<uint4 *>(out)[0] = make_uint4(h[0][d], h[1][d], h[2][d], h[3][d]);
            reinterpret_cast<uint4 *>(out)[1] = make_uint4(h[4][d], h[5][d], h[6][d], h[7][d]);
        }
        else
            for (int w = 0; w < kUpRows; ++w)
                if (r0 + w < ranges) out[w] = h[w][d];
    }
}

// Block digit blockIdx.x: exclusive prefix of the digit's count over the
// ranges (a range's count is the sum of its `parts` rows) and the total.
// Each thread takes 16 consecutive ranges (4096 per block and round, with a
// running carry between rounds), so a round costs one block scan.
__global__ __launch_bounds__(kSortThreads) void scanDigitKernel(const std::uint32_t *__restrict__ hist,
                                                                std::size_t ranges, int parts,
                                                                std::uint32_t *rangeOff, std::uint32_t *digitTotal)
{
    constexpr int kPer = 16;
    __shared__ std::uint32_t waveTotals[2][kWaves];
    const std::size_t d = blockIdx.x;
    const std::size_t rows = ranges * std::size_t(parts);
    const std::uint32_t *h = hist + d * rows;
    std::uint32_t *out = rangeOff + d * ranges;
    // 16-byte vector access when every digit's row and this block's rows are aligned
    const bool vec = parts == 1 && ranges % 4 == 0 && (reinterpret_cast<std::uintptr_t>(hist) & 15) == 0 &&
                     (reinterpret_cast<std::uintptr_t>(rangeOff) & 15) == 0;
    const unsigned lane = laneId();
    const int wave = threadIdx.x >> 6;
    std::uint32_t carry = 0;
    int buf = 0;
    for (std::size_t r0 = 0; r0 < ranges; r0 += std::size_t(kPer) * kSortThreads)
    {
        const std::size_t first = r0 + std::size_t(threadIdx.x) * kPer;
        std::uint32_t v[kPer];
        if (vec && first + kPer <= ranges)
        {
#pragma unroll
            for (int q = 0; q < kPer / 4; ++q)
            {
                const uint4 c = reinterpret_cast<const uint4 *>(h + first)[q];
                v[4 * q] = c.x, v[4 * q + 1] = c.y, v[4 * q + 2] = c.z, v[4 * q + 3] = c.w;
            }
        }
        else
        {
#pragma unroll
            for (int k = 0; k < kPer; ++k)
            {
                v[k] = 0;
                if (first + k < ranges)
                    for (int q = 0; q < parts; ++q) v[k] += h[(first + k) * unsigned(parts) + unsigned(q)];
            }
        }
        std::uint32_t sum = 0; // v becomes the thread-exclusive prefix
#pragma unroll
        for (int k = 0; k < kPer; ++k)
        {
            const std::uint32_t t = v[k];
            v[k] = sum;
            sum += t;
        }
        std::uint32_t inc = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const std::uint32_t t = __shfl_up(inc, o, 64);
            if (lane >= unsigned(o)) inc += t;
        }
        if (lane == 63) waveTotals[buf][wave] = inc;
        __syncthreads(); // double-buffered totals: one barrier per round
        std::uint32_t before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w)
        {
            const std::uint32_t t = waveTotals[buf][w];
            before += w < wave ? t : 0u;
            total += t;
        }
        const std::uint32_t base = carry + before + inc - sum;
        if (vec && first + kPer <= ranges)
        {
#pragma unroll
            for (int q = 0; q < kPer / 4; ++q)
                reinterpret_cast<uint4 *>(out + first)[q] =
                    make_uint4(base + v[4 * q], base + v[4 * q + 1], base + v[4 * q + 2], base + v[4 * q + 3]);
        }
        else
        {
#pragma unroll
            for (int k = 0; k < kPer; ++k)
                if (first + k < ranges) out[first + k] = base + v[k];
        }
        carry += total;
        buf ^= 1;
    }
    if (threadIdx.x == 0) digitTotal[d] = carry;
}

// Bit B of v as 0 or ~0 (v_bfe_i32). The extract is kept opaque so that the
// ballot below is one compare of this value with zero: the compiler would
// otherwise test the bit a second time with a shift and a sign compare.
template <int B>
__device__ inline std::uint32_t bitAsMask(unsigned v)
{
    std::uint32_t m;
    asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(m) : "v"(v), "n"(B));
    return m;
}

// Lanes of the wave whose BITS-bit digit equals this lane's: for every digit
// bit, the ballot of lanes with the bit set, complemented where this lane's
// bit is clear (a bit equal on every lane leaves the mask unchanged). With
// mask = 0 or ~0, each step is one three-input bitwise operation (v_bitop3)
// per 32-bit half; the steps carry no branches.
template <int B, int BITS>
__device__ inline void laneMatchStep(unsigned dg, std::uint32_t &lo, std::uint32_t &hi)
{
    if constexpr (B < BITS)
    {
        const std::uint32_t mask = bitAsMask<B>(dg);
        const std::uint64_t ballot = __ballot(int(mask));
        // x & ~(ballot ^ mask): truth table 0x90 over (x, ballot, mask)
        lo = __builtin_amdgcn_bitop3_b32(lo, std::uint32_t(ballot), mask, 0x90);
        hi = __builtin_amdgcn_bitop3_b32(hi, std::uint32_t(ballot >> 32), mask, 0x90);
        laneMatchStep<B + 1, BITS>(dg, lo, hi);
    }
}
template <int BITS>
__device__ inline std::uint64_t laneMatchBits(unsigned dg)
{
    std::uint32_t lo = ~0u, hi = ~0u;
    laneMatchStep<0, BITS>(dg, lo, hi);
    return (std::uint64_t(hi) << 32) | lo;
}
__device__ inline std::uint64_t laneMatch(unsigned dg) { return laneMatchBits<kRadixBits>(dg); }

// Number of lanes below this one in `mask` (v_mbcnt_lo / v_mbcnt_hi).
__device__ inline std::uint32_t lanesBelow(std::uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(std::uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(std::uint32_t(mask), 0u));
}

// Exclusive scan of the 256 per-digit values held by threads 0..255 of a
// block of any size (all threads must call it).
__device__ inline std::uint32_t digitExclusiveScan(std::uint32_t v, std::uint32_t *waveTotals)
{
    const unsigned lane = laneId();
    const int wave = threadIdx.x >> 6;
    std::uint32_t inc = v;
    if (wave < kRadix / 64)
    {
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const std::uint32_t t = __shfl_up(inc, o, 64);
            if (lane >= unsigned(o)) inc += t;
        }
        if (lane == 63) waveTotals[wave] = inc;
    }
    __syncthreads();
    std::uint32_t before = 0;
#pragma unroll
    for (int w = 0; w < kRadix / 64; ++w)
        if (w < wave) before += waveTotals[w];
    __syncthreads();
    return before + inc - v;
}

// Second launch bound = minimum waves per SIMD: two blocks of 512 x 8 (or
// one of 1024) per CU, i.e. at most 128 VGPRs; larger 512-thread tiles fit
// once per CU (LDS), as do 256-thread blocks' register budgets of 256.
//
// kSeg: segmented sort. The range comes from segRanges[blockIdx.x] and
// rangeOff holds its absolute per-digit start ([range][digit]); values of the
// first pass are slots within the segment.
//
// kPre: the keys (and values) of the next tile are loaded while the current
// tile is ranked and written (register double buffering).
//
// kMsd: a partition level of the MSD sort. The block loops over the
// *msd.rangeCount ranges of segRanges (grid-stride), each of at most one tile
// (the level plans ranges of THREADS * ITEMS keys); a range's digit d starts
// at its segment's begin + the segment's keys of smaller digits (an exclusive
// scan of msd.subTotals[seg]) + msd.off[d][range].
struct MsdScatterArgs
{
    const std::uint32_t *rangeCount = nullptr;
    const MsdSeg *segs = nullptr;
    const std::uint32_t *subTotals = nullptr; // [segment][digit] keys of each digit
    const std::uint32_t *off = nullptr;       // [digit][stride] offset of a range within its segment
    std::size_t stride = 0;
};

template <int THREADS, int ITEMS, bool kVals, bool kIota, bool kDigits, bool kSeg = false, bool kPre = false,
          bool kMsd = false>
__global__ __launch_bounds__(THREADS, (THREADS == 512 && ITEMS > 8) || THREADS < 512 ? 2 : 4) void scatterKernel(
    const Key *__restrict__ kin, Key *__restrict__ kout, const std::uint32_t *__restrict__ vin,
    std::uint32_t *__restrict__ vout, std::uint8_t *__restrict__ digitsOut, std::size_t n, std::size_t rangeKeys,
    const std::uint32_t *__restrict__ rangeOff, const std::uint32_t *__restrict__ digitTotal, std::size_t ranges,
    int shift, int nextShift, const SegRange *__restrict__ segRanges = nullptr, MsdScatterArgs msd = {})
{
    static_assert(!(kMsd && kPre), "an MSD level's tile loads its own keys");
    constexpr int kW = THREADS / 64;
    constexpr int kTile = ITEMS * THREADS;
    __shared__ Key stageKeys[kTile];                         // the tile in output order
    __shared__ std::uint32_t stageVals[kVals ? kTile : 1];
    __shared__ std::uint32_t waveCount[kW][kRadix]; // per-wave digit counts, then offsets
    __shared__ std::uint32_t sink[kW][64];          // ranking adds of non-leader lanes
    __shared__ std::uint32_t running[kRadix];
    __shared__ std::uint32_t adjust[kRadix];
    __shared__ std::uint32_t scratch[kRadix / 64];

    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const unsigned lane = laneId();
    auto body = [&](const std::size_t r) {
    std::size_t rBegin;
    std::uint32_t rLen, valueBase; // range-relative 32-bit offsets (n < 2^32)
    std::uint32_t seg = 0;
    if constexpr (kMsd || kSeg)
    {
        const SegRange d = segRanges[r];
        rBegin = d.begin;
        rLen = d.len;
        valueBase = d.localBase;
        seg = d.seg;
    }
    else
    {
        rBegin = r * rangeKeys;
        rLen = std::uint32_t((rBegin + rangeKeys < n ? rBegin + rangeKeys : n) - rBegin);
        valueBase = std::uint32_t(rBegin);
    }
    if (rLen == 0) return; // block-uniform
    const Key *kr = kin + rBegin;
    const std::uint32_t *vr = vin + rBegin;
    const std::uint32_t last = rLen - 1;
    constexpr bool kPreVals = kVals && !kIota;
    Key preKey[ITEMS];
    std::uint32_t preVal[kPreVals ? ITEMS : 1];
    // Loads and stores are kept free of per-lane branches (clamped indices
    // and selects; the partial last tile is a separate copy of the body) so
    // that the wait for a tile's keys counts only the loads, not the stores
    // still in flight from the previous tile.
    auto prefetch = [&](std::uint32_t first) {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
        {
            const std::uint32_t idx = first + std::uint32_t(j) * 64;
            const std::uint32_t c = idx < rLen ? idx : last;
            const Key k = kr[c];
            preKey[j] = idx < rLen ? k : kInvalidKey;
            if constexpr (kPreVals) preVal[j] = vr[c];
        }
    };
    // The range's digit totals and offsets are loaded first, then the first
    // tile's keys; both are in flight together, and the wait for the former
    // (loads complete in order) leaves the tile's loads outstanding during the
    // digit scan. (An MSD level's tile loads its own keys after the scan: the
    // preloaded form of that kernel needs more than 128 VGPRs and spills.)
    if constexpr (kMsd)
    {
        const std::uint32_t total = tid < kRadix ? msd.subTotals[std::size_t(seg) * kRadix + tid] : 0u;
        const std::uint32_t off = tid < kRadix ? msd.off[tid * msd.stride + r] : 0u;
        const std::uint32_t digitBase = std::uint32_t(msd.segs[seg].begin) + digitExclusiveScan(total, scratch);
        if (tid < kRadix) running[tid] = digitBase + off;
    }
    else if constexpr (kSeg)
    {
        const std::uint32_t off = tid < kRadix ? rangeOff[r * kRadix + tid] : 0u;
        prefetch(std::uint32_t(wave) * (ITEMS * 64) + lane);
        if (tid < kRadix) running[tid] = off;
    }
    else
    {
        const std::uint32_t total = tid < kRadix ? digitTotal[tid] : 0u;
        const std::uint32_t off = tid < kRadix ? rangeOff[std::size_t(tid) * ranges + r] : 0u;
        prefetch(std::uint32_t(wave) * (ITEMS * 64) + lane);
        const std::uint32_t digitBase = digitExclusiveScan(total, scratch);
        if (tid < kRadix) running[tid] = digitBase + off;
    }
    for (int i = tid; i < kW * kRadix; i += THREADS) (&waveCount[0][0])[i] = 0;
    __syncthreads();

    // One tile: 5 barriers (ranked, wave totals, offsets, staged, written).
    // The counters are cleared during the write-out, ready for the next tile.
    auto tile = [&](std::uint32_t base, auto fullTag, auto preTag) {
        constexpr bool kFull = decltype(fullTag)::value;
        constexpr bool kFromPre = decltype(preTag)::value; // keys already in preKey
        constexpr bool kLoadVals = kVals && !kIota;
        // ---- load keys (wave-striped: item j of lane l is element j*64 + l of the wave's chunk)
        Key key[ITEMS];
        std::uint32_t val[kLoadVals ? ITEMS : 1];
        const std::uint32_t waveFirst = base + std::uint32_t(wave) * (ITEMS * 64) + lane;
        if constexpr (kPre)
        {
#pragma unroll
            for (int j = 0; j < ITEMS; ++j)
            {
                key[j] = preKey[j];
                if constexpr (kPreVals) val[j] = preVal[j];
            }
            if (kFull) prefetch(waveFirst + kTile); // the next tile, or clamped reloads of the last key
        }
        else if constexpr (kFromPre)
        {
#pragma unroll
            for (int j = 0; j < ITEMS; ++j)
            {
                key[j] = preKey[j];
                if constexpr (kPreVals) val[j] = preVal[j];
            }
        }
        else
        {
#pragma unroll
            for (int j = 0; j < ITEMS; ++j)
            {
                const std::uint32_t idx = waveFirst + std::uint32_t(j) * 64;
                const std::uint32_t c = kFull || idx < rLen ? idx : last;
                const Key k = kr[c];
                key[j] = kFull || idx < rLen ? k : kInvalidKey; // padding ranks last with digit 255
                if constexpr (kLoadVals) val[j] = vr[c];        // padding values are never written
            }
        }

        // ---- rank within the wave: lanes with equal digits via 8 ballots; the
        // lowest such lane adds the group's size to the wave's digit counter
        // (the other lanes add 0 to a slot of their own, so every lane issues
        // the same LDS atomic and the adds of all items are in flight at
        // once; one wave's LDS operations execute in order), then the old
        // count is read from that lane.
        std::uint32_t rank[ITEMS], old[ITEMS];
        int leader[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
        {
            const unsigned dg = unsigned(key[j] >> shift) & (kRadix - 1);
            const std::uint64_t match = laneMatch(dg);
            const std::uint32_t lower = lanesBelow(match);
            const bool lead = lower == 0;
            std::uint32_t *slot = lead ? &waveCount[wave][dg] : &sink[wave][lane];
            old[j] = atomicAdd(slot, lead ? std::uint32_t(__popcll(match)) : 0u);
            leader[j] = __builtin_ctzll(match);
            rank[j] = lower;
        }
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) rank[j] += std::uint32_t(__shfl(int(old[j]), leader[j], 64));
        __syncthreads();

        // ---- per digit: wave offsets, tile count, tile-local digit start
        // (threads 0..255, one digit each; wave-inclusive scan, then the
        // totals of the four digit waves)
        std::uint32_t tileCount = 0, inc = 0;
        if (tid < kRadix)
        {
#pragma unroll
            for (int w = 0; w < kW; ++w)
            {
                const std::uint32_t c = waveCount[w][tid];
                waveCount[w][tid] = tileCount;
                tileCount += c;
            }
            inc = tileCount;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1)
            {
                const std::uint32_t t = __shfl_up(inc, o, 64);
                if (lane >= unsigned(o)) inc += t;
            }
            if (lane == 63) scratch[wave] = inc;
        }
        __syncthreads();
        if (tid < kRadix)
        {
            std::uint32_t before = 0;
#pragma unroll
            for (int w = 0; w < kRadix / 64; ++w)
                if (w < wave) before += scratch[w];
            const std::uint32_t tileStart = before + inc - tileCount;
#pragma unroll
            for (int w = 0; w < kW; ++w) waveCount[w][tid] += tileStart;
            adjust[tid] = running[tid] - tileStart;
            running[tid] += tileCount;
        }
        __syncthreads();

        // ---- stage keys (and values) in output order
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
        {
            const std::uint32_t pos = rank[j] + waveCount[wave][unsigned(key[j] >> shift) & (kRadix - 1)];
            stageKeys[pos] = key[j];
            if constexpr (kVals)
            {
                std::uint32_t v;
                if constexpr (kIota) v = valueBase + waveFirst + std::uint32_t(j) * 64;
                else v = val[j];
                stageVals[pos] = v;
            }
        }
        __syncthreads();

        // ---- write runs of equal digits; clear the counters for the next tile
        for (int i = tid; i < kW * kRadix; i += THREADS) (&waveCount[0][0])[i] = 0;
        const int valid = kFull ? kTile : int(rLen - base);
#pragma unroll
        for (int q = 0; q < ITEMS; ++q)
        {
            const int i = q * THREADS + tid;
            if (kFull || i < valid)
            {
                const Key k = stageKeys[i];
                const std::uint32_t dst = adjust[unsigned(k >> shift) & (kRadix - 1)] + std::uint32_t(i);
                kout[dst] = k;
                if (kDigits) digitsOut[dst] = std::uint8_t(k >> nextShift);
                if (kVals) vout[dst] = stageVals[i];
            }
        }
        __syncthreads();
    };

    std::uint32_t base = 0;
    if constexpr (kMsd) // one tile
    {
        if (rLen >= std::uint32_t(kTile)) tile(0u, std::true_type{}, std::false_type{});
        else tile(0u, std::false_type{}, std::false_type{});
    }
    else if constexpr (kPre)
    {
        for (; rLen - base >= std::uint32_t(kTile); base += kTile) tile(base, std::true_type{}, std::true_type{});
        if (base < rLen) tile(base, std::false_type{}, std::true_type{});
    }
    else
    {
        // the first tile from the registers loaded before the prologue (a
        // separate copy, so that they are dead in the loop)
        if (rLen >= std::uint32_t(kTile))
        {
            tile(0u, std::true_type{}, std::true_type{});
            base = kTile;
        }
        else
        {
            tile(0u, std::false_type{}, std::true_type{});
            base = rLen;
        }
        for (; rLen - base >= std::uint32_t(kTile); base += kTile) tile(base, std::true_type{}, std::false_type{});
        if (base < rLen) tile(base, std::false_type{}, std::false_type{});
    }
    }; // body: every path ends with a barrier, so the next range may reuse the shared state

    if constexpr (kMsd)
    {
        const std::uint32_t count = *msd.rangeCount < msd.stride ? *msd.rangeCount : std::uint32_t(msd.stride);
        for (std::uint32_t w = blockIdx.x; w < count; w += gridDim.x) body(xcdRow(w, count));
    }
    else
        body(xcdRow(blockIdx.x, ranges)); // gridDim.x == ranges
}

// ---- wide digits: LSD passes of 11 bits ------------------------------------------
// The digit width the autotuner may pick for LSD sorts (below the MSD
// threshold and kWideMaxKeys keys): 6 passes of 11 bits instead of 8 of 8 cover the 64-bit key, 18
// launches instead of 24. One tile of 2048 keys per range; per pass a
// histogram of each range (digit-major rows), one scan block per digit over
// the ranges, and a scatter that ranks each key within its wave by 11
// ballots, stages the tile in output order in LDS and writes it out.
// Stable, like the 8-bit passes: ties keep their input order.
constexpr int kWideBits = 11;
constexpr int kWideRadix = 1 << kWideBits;
constexpr int kWideThreads = 256, kWideItems = 8;
constexpr int kWideTile = kWideThreads * kWideItems;
constexpr int kWidePasses = (64 + kWideBits - 1) / kWideBits; // even: the result ends where it started
constexpr int kWideDigitsPerThread = kWideRadix / kWideThreads;
static_assert(kWidePasses % 2 == 0, "an even pass count leaves the result in the input buffers");

// Exclusive scan of one value per thread of a kWideThreads block; total =
// the sum over the block. All threads must call it.
__device__ inline std::uint32_t wideBlockScan(std::uint32_t v, std::uint32_t *waveTotals, std::uint32_t &total)
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
    for (int w = 0; w < kWideThreads / 64; ++w)
    {
        if (w < wave) before += waveTotals[w];
        total += waveTotals[w];
    }
    __syncthreads();
    return before + inc - v;
}

// hist[d * ranges + r] = keys of range r (keys [r * kWideTile, ...)) whose
// digit at `shift` is d.
__global__ __launch_bounds__(kWideThreads) void wideUpsweepKernel(const Key *__restrict__ keys, std::size_t n,
                                                                  int shift, std::uint32_t *__restrict__ hist,
                                                                  std::uint32_t ranges)
{
    __shared__ std::uint32_t h[kWideRadix];
    for (int d = threadIdx.x; d < kWideRadix; d += kWideThreads) h[d] = 0;
    __syncthreads();
    const std::size_t base = std::size_t(blockIdx.x) * kWideTile;
#pragma unroll
    for (int j = 0; j < kWideItems; ++j)
    {
        const std::size_t i = base + std::size_t(j) * kWideThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[unsigned(keys[i] >> shift) & (kWideRadix - 1)], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kWideRadix; d += kWideThreads) hist[std::size_t(d) * ranges + blockIdx.x] = h[d];
}

// Block d: exclusive prefix of digit d's counts over the ranges, and its total.
__global__ __launch_bounds__(kWideThreads) void wideScanKernel(const std::uint32_t *__restrict__ hist,
                                                               std::uint32_t *__restrict__ off,
                                                               std::uint32_t *__restrict__ total,
                                                               std::uint32_t ranges)
{
    __shared__ std::uint32_t wt[kWideThreads / 64];
    const std::size_t row = std::size_t(blockIdx.x) * ranges;
    std::uint32_t carry = 0;
    for (std::uint32_t b = 0; b < ranges; b += kWideThreads)
    {
        const std::uint32_t r = b + threadIdx.x;
        const std::uint32_t v = r < ranges ? hist[row + r] : 0u;
        std::uint32_t all;
        const std::uint32_t ex = wideBlockScan(v, wt, all);
        if (r < ranges) off[row + r] = carry + ex;
        carry += all;
    }
    if (threadIdx.x == 0) total[blockIdx.x] = carry;
}

// Block r scatters range r (one tile). kIota: the values are the keys'
// input positions (the first pass of a sort by slot).
template <bool kVals, bool kIota>
__global__ __launch_bounds__(kWideThreads) void wideScatterKernel(const Key *__restrict__ kin, Key *__restrict__ kout,
                                                                  const std::uint32_t *__restrict__ vin,
                                                                  std::uint32_t *__restrict__ vout, std::size_t n,
                                                                  int shift, const std::uint32_t *__restrict__ off,
                                                                  const std::uint32_t *__restrict__ total,
                                                                  std::uint32_t ranges)
{
    constexpr int kW = kWideThreads / 64;
    constexpr int kD = kWideDigitsPerThread;
    constexpr bool kLoadVals = kVals && !kIota;
    __shared__ Key stageKeys[kWideTile];
    __shared__ std::uint32_t stageVals[kVals ? kWideTile : 1];
    __shared__ std::uint32_t waveCount[kW][kWideRadix]; // per-wave digit counts, then tile positions
    __shared__ std::uint32_t sink[kW][64];              // ranking adds of non-leader lanes
    __shared__ std::uint32_t adjust[kWideRadix];        // output index of a staged position, per digit
    __shared__ std::uint32_t wt[kW];
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const unsigned lane = laneId();
    const std::uint32_t r = blockIdx.x;
    const std::size_t rBegin = std::size_t(r) * kWideTile;
    const std::uint32_t rLen = std::uint32_t(n - rBegin < std::size_t(kWideTile) ? n - rBegin : kWideTile);

    // the range's start of every digit (thread t: digits kD t .. kD t + kD - 1)
    std::uint32_t tot[kD], sum = 0;
#pragma unroll
    for (int q = 0; q < kD; ++q)
    {
        tot[q] = total[kD * tid + q];
        sum += tot[q];
    }
    std::uint32_t all;
    std::uint32_t run = wideBlockScan(sum, wt, all);
#pragma unroll
    for (int q = 0; q < kD; ++q)
    {
        const int d = kD * tid + q;
        adjust[d] = run + off[std::size_t(d) * ranges + r];
        run += tot[q];
    }
    for (int i = tid; i < kW * kWideRadix; i += kWideThreads) (&waveCount[0][0])[i] = 0;

    // keys (wave-striped: item j of lane l is element j * 64 + l of the
    // wave's 512); padding ranks last with the largest digit
    Key key[kWideItems];
    std::uint32_t val[kLoadVals ? kWideItems : 1];
    const std::uint32_t waveFirst = std::uint32_t(wave) * (kWideItems * 64) + lane;
#pragma unroll
    for (int j = 0; j < kWideItems; ++j)
    {
        const std::uint32_t idx = waveFirst + std::uint32_t(j) * 64;
        const std::uint32_t c = idx < rLen ? idx : rLen - 1;
        const Key k = kin[rBegin + c];
        key[j] = idx < rLen ? k : kInvalidKey;
        if constexpr (kLoadVals) val[j] = vin[rBegin + c];
    }
    __syncthreads();

    // rank within the wave: the lowest lane of each digit group adds the
    // group's size to the wave's counter (the others add 0 to a slot of
    // their own; one wave's LDS operations execute in order)
    std::uint32_t rank[kWideItems], old[kWideItems];
    int leader[kWideItems];
#pragma unroll
    for (int j = 0; j < kWideItems; ++j)
    {
        const unsigned dg = unsigned(key[j] >> shift) & (kWideRadix - 1);
        const std::uint64_t match = laneMatchBits<kWideBits>(dg);
        const std::uint32_t lower = lanesBelow(match);
        const bool lead = lower == 0;
        std::uint32_t *slot = lead ? &waveCount[wave][dg] : &sink[wave][lane];
        old[j] = atomicAdd(slot, lead ? std::uint32_t(__popcll(match)) : 0u);
        leader[j] = __builtin_ctzll(match);
        rank[j] = lower;
    }
#pragma unroll
    for (int j = 0; j < kWideItems; ++j) rank[j] += std::uint32_t(__shfl(int(old[j]), leader[j], 64));
    __syncthreads();

    // per digit: wave offsets and the tile count, then the digits' starts in the tile
    std::uint32_t cnt[kD], tsum = 0;
#pragma unroll
    for (int q = 0; q < kD; ++q)
    {
        const int d = kD * tid + q;
        std::uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < kW; ++w)
        {
            const std::uint32_t x = waveCount[w][d];
            waveCount[w][d] = c;
            c += x;
        }
        cnt[q] = c;
        tsum += c;
    }
    std::uint32_t tstart = wideBlockScan(tsum, wt, all);
#pragma unroll
    for (int q = 0; q < kD; ++q)
    {
        const int d = kD * tid + q;
#pragma unroll
        for (int w = 0; w < kW; ++w) waveCount[w][d] += tstart;
        adjust[d] -= tstart;
        tstart += cnt[q];
    }
    __syncthreads();

    // stage in output order, then write
#pragma unroll
    for (int j = 0; j < kWideItems; ++j)
    {
        const std::uint32_t pos = rank[j] + waveCount[wave][unsigned(key[j] >> shift) & (kWideRadix - 1)];
        stageKeys[pos] = key[j];
        if constexpr (kVals)
        {
            std::uint32_t v;
            if constexpr (kIota) v = std::uint32_t(rBegin) + waveFirst + std::uint32_t(j) * 64;
            else v = val[j];
            stageVals[pos] = v;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kWideItems; ++q)
    {
        const int i = q * kWideThreads + tid;
        if (std::uint32_t(i) < rLen)
        {
            const Key k = stageKeys[i];
            const std::uint32_t dst = adjust[unsigned(k >> shift) & (kWideRadix - 1)] + std::uint32_t(i);
            kout[dst] = k;
            if constexpr (kVals) vout[dst] = stageVals[i];
        }
    }
}

// ---- segmented sort kernels --------------------------------------------------
constexpr int kSegThreads = 512, kSegItems = 8; // tile shape of the segmented scatter

template <class T, KeyKind K>
__global__ __launch_bounds__(kSortThreads) void segKeygenKernel(const KeygenParams<T> p, const SegCoords<T> *coords,
                                                                const SegRange *ranges, Key *keys,
                                                                std::uint32_t *hist, int histShift)
{
    __shared__ std::uint32_t h[kRadix];
    __shared__ std::uint16_t henc[K == KeyKind::Hilbert ? kHilbertRows : 1];
    h[threadIdx.x] = 0;
    if constexpr (K == KeyKind::Hilbert) stageHilbertEnc(henc);
    __syncthreads();
    const SegRange d = ranges[blockIdx.x];
    const SegCoords<T> c = coords[d.seg];
    // p stays unmodified (in the kernel arguments): the binary key's
    // per-depth axis table is then indexed there, not in a private copy
    for (std::uint32_t j = threadIdx.x; j < d.len; j += kSortThreads)
    {
        const std::size_t i = std::size_t(d.localBase) + j;
        const Key k = keyOf<T, K, false>(p, nullptr, henc, c.x[i], c.y[i], c.z[i]);
        keys[d.begin + j] = k;
        countDigit(h, unsigned(k >> histShift) & (kRadix - 1));
    }
    __syncthreads();
    hist[std::size_t(blockIdx.x) * kRadix + threadIdx.x] = h[threadIdx.x];
}

// Digit histogram of one range from the digit bytes, [range][digit] layout.
__global__ __launch_bounds__(kSortThreads) void segUpsweepDigitsKernel(const std::uint8_t *__restrict__ digits,
                                                                       const SegRange *ranges, std::uint32_t *hist)
{
    __shared__ std::uint32_t h[kWaves][kRadix];
#pragma unroll
    for (int w = 0; w < kWaves; ++w) h[w][threadIdx.x] = 0;
    __syncthreads();
    std::uint32_t *mine = h[threadIdx.x >> 6];
    const SegRange d = ranges[blockIdx.x];
    const std::uint8_t *src = digits + d.begin;
    for (std::uint32_t j = threadIdx.x; j < d.len; j += kSortThreads) atomicAdd(&mine[src[j]], 1u);
    __syncthreads();
    std::uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) t += h[w][threadIdx.x];
    hist[std::size_t(blockIdx.x) * kRadix + threadIdx.x] = t;
}

// One block per segment, one thread per digit: the absolute start of every
// (range, digit) of the segment = segment begin + keys of smaller digits in
// the segment + keys of this digit in earlier ranges of the segment.
__global__ __launch_bounds__(kSortThreads) void segScanKernel(const std::uint32_t *__restrict__ hist,
                                                              const std::uint32_t *segRange,
                                                              const std::uint64_t *segBegin, std::uint32_t *off,
                                                              std::uint32_t *segTotals)
{
    __shared__ std::uint32_t waveTotals[kWaves];
    const unsigned d = threadIdx.x;
    const std::uint32_t r0 = segRange[blockIdx.x], r1 = segRange[blockIdx.x + 1];
    // eight ranges' counts in flight at a time (a member of 2M keys has ~500 ranges)
    constexpr std::uint32_t kAhead = 8;
    std::uint32_t run = 0;
    for (std::uint32_t r = r0; r < r1; r += kAhead)
    {
        std::uint32_t c[kAhead];
#pragma unroll
        for (std::uint32_t u = 0; u < kAhead; ++u) c[u] = r + u < r1 ? hist[std::size_t(r + u) * kRadix + d] : 0u;
#pragma unroll
        for (std::uint32_t u = 0; u < kAhead; ++u)
            if (r + u < r1)
            {
                off[std::size_t(r + u) * kRadix + d] = run;
                run += c[u];
            }
    }
    if (segTotals) segTotals[std::size_t(blockIdx.x) * kRadix + d] = run; // the segment's count of digit d
    const std::uint32_t base = std::uint32_t(segBegin[blockIdx.x]) + blockExclusiveScan(run, waveTotals);
    for (std::uint32_t r = r0; r < r1; r += kAhead)
    {
        std::uint32_t c[kAhead];
#pragma unroll
        for (std::uint32_t u = 0; u < kAhead; ++u) c[u] = r + u < r1 ? off[std::size_t(r + u) * kRadix + d] : 0u;
#pragma unroll
        for (std::uint32_t u = 0; u < kAhead; ++u)
            if (r + u < r1) off[std::size_t(r + u) * kRadix + d] = c[u] + base;
    }
}

// ---- MSD sort ----------------------------------------------------------------
// Level 1 partitions all keys by bits 56..63 (the pass machinery above);
// level L >= 2 partitions only the buckets of more than kMsdChunk keys left
// by level L-1, by bits 64-8L .. 71-8L. Every bucket of at most kMsdChunk
// keys is final: runs of them become chunks, sorted in LDS on the bits below
// their parent's digit. All lists and counts live on the device, so the
// whole sort is one fixed sequence of launches (graph-capturable); kernels
// loop over the device-side counts. Levels alternate output buffers: level L
// writes the odd-level buffer (Y) when L is odd, else X.
constexpr int kLocThreads = 256, kLocItems = 8;       // LDS chunk sort block shape
constexpr std::uint32_t kMsdChunk = kLocThreads * kLocItems; // LDS chunk capacity (keys)
constexpr int kMsdThreads = 512, kMsdItems = 8;       // scatter tile of levels >= 2
constexpr std::uint32_t kMsdTile = kMsdThreads * kMsdItems;
constexpr unsigned kMsdScanRows = 16;                 // chunk rows of the level scan's grid (x 256 digits)
constexpr unsigned kRootPlanBlocks = 32;              // blocks sharing the range writes of the root plan

__global__ void msdResetKernel(MsdCounters *c, MsdSeg *root, std::size_t n)
{
    c->segs[0] = c->segs[1] = 0;
    c->ranges[0] = c->ranges[1] = 0;
    c->chunks = 0;
    c->overflow = 0;
    *root = MsdSeg{0, std::uint32_t(n), 0, 0, 0};
}

// One block per parent segment (grid-stride over *parentCount): thread d
// takes sub-bucket d (size totals[parent][d]). Sub-buckets of more than
// kMsdChunk keys become segments of the next level; the block writes their
// tile ranges together (thread t takes ranges t, t+256, ...). The others are
// grouped, in order, into chunks of at most kMsdChunk keys: a chunk starting
// at sub-bucket d extends up to the next big sub-bucket or until one more
// would exceed kMsdChunk keys (the same cuts as a sequential greedy pass);
// each thread finds that end by a binary search and thread 0 follows the
// chunk starts. At the last level every sub-bucket holds equal keys and the
// parent becomes one chunk that needs no sorting. A parent reserves its
// entries of the three lists in one round trip (three atomics issued
// together). Root plan (parentCount == nullptr: the whole array, lists
// empty): every block plans the root, block 0 sets the list counts and
// writes the segments and chunks, and the blocks share the range writes.
__global__ __launch_bounds__(kSortThreads) void msdPlanKernel(
    int level, const MsdSeg *__restrict__ parents, const std::uint32_t *parentCount,
    const std::uint32_t *__restrict__ totals, MsdCounters *cnt, int nextSet, MsdSeg *nextSegs,
    SegRange *nextRanges, std::uint8_t *nextFirst, std::uint32_t maxSegs, std::uint32_t maxRanges, MsdChunk *chunks,
    std::uint32_t maxChunks, std::uint32_t maxParents)
{
    __shared__ std::uint32_t waveTotals[kWaves];
    __shared__ std::uint32_t relS[kRadix + 1];    // sub-bucket start within the parent
    __shared__ std::uint32_t inclSmall[kRadix];   // prefix of the small sizes, inclusive
    __shared__ std::uint16_t nextBig[kRadix + 1]; // first big sub-bucket at or after d
    __shared__ std::uint16_t nextStart[kRadix + 1]; // first nonempty small sub-bucket at or after d
    __shared__ std::uint16_t bigAt[kRadix + 1], smallAt[kRadix + 1];
    __shared__ std::uint16_t chunkEnd[kRadix];
    __shared__ std::uint32_t rangeRankOf[kRadix + 1]; // ranges before the k-th big sub-bucket
    __shared__ std::uint16_t bigIndex[kRadix];        // sub-bucket of the k-th big one
    __shared__ std::uint32_t chBegin[kRadix], chLen[kRadix];
    __shared__ std::uint32_t base[3];
    __shared__ std::uint32_t nChunks, nBig, nRanges;
    const unsigned d = threadIdx.x;
    const bool root = parentCount == nullptr;
    const std::uint32_t count = root ? 1u : (*parentCount < maxParents ? *parentCount : maxParents);
    const bool last = level >= kSortPasses;
    const std::uint8_t parity = std::uint8_t(level & 1);
    // a chunk spans sub-buckets of one parent: sort on the bits below the parent's digit
    const std::uint8_t bits = last ? 0 : std::uint8_t(level == 1 ? 64 : 72 - 8 * level);
    const bool lead = !root || blockIdx.x == 0; // writes counts, segments and chunks
    for (std::uint32_t p = root ? 0u : blockIdx.x; p < count; p += root ? count : gridDim.x)
    {
        const MsdSeg par = parents[p];
        const std::uint32_t s = totals[std::size_t(p) * kRadix + d];
        const std::uint32_t rel = blockExclusiveScan(s, waveTotals);
        const bool isBig = !last && s > kMsdChunk;
        const bool isSmall = !last && !isBig && s > 0;
        const std::uint32_t nr = isBig ? (s + kMsdTile - 1) / kMsdTile : 0u;
        const std::uint32_t bigRank = blockExclusiveScan(isBig ? 1u : 0u, waveTotals);
        const std::uint32_t rangeRank = blockExclusiveScan(nr, waveTotals);
        const std::uint32_t smallRank = blockExclusiveScan(isSmall ? 1u : 0u, waveTotals);
        const std::uint32_t smallBefore = blockExclusiveScan(isSmall ? s : 0u, waveTotals);
        relS[d] = rel;
        inclSmall[d] = smallBefore + (isSmall ? s : 0u);
        if (isBig)
        {
            bigAt[bigRank] = std::uint16_t(d);
            bigIndex[bigRank] = std::uint16_t(d);
            rangeRankOf[bigRank] = rangeRank;
        }
        if (isSmall) smallAt[smallRank] = std::uint16_t(d);
        if (d == kRadix - 1)
        {
            relS[kRadix] = rel + s;
            const std::uint32_t nb = bigRank + (isBig ? 1u : 0u), ns = smallRank + (isSmall ? 1u : 0u);
            const std::uint32_t nrr = rangeRank + nr;
            bigAt[nb] = kRadix;
            smallAt[ns] = kRadix;
            rangeRankOf[nb] = nrr;
            nBig = nb;
            nRanges = nrr;
        }
        __syncthreads();
        nextBig[d] = bigAt[bigRank];     // d itself when big
        nextStart[d] = smallAt[smallRank]; // d itself when small
        if (d == 0) nextBig[kRadix] = kRadix, nextStart[kRadix] = kRadix;
        __syncthreads();
        // end of a chunk starting at small sub-bucket d
        if (isSmall)
        {
            const std::uint32_t limit = inclSmall[d] - s + kMsdChunk; // small keys before d + capacity
            const unsigned stop = nextBig[d];
            unsigned lo = d, hi = stop; // inclSmall[lo] <= limit; find the last such index below stop
            while (hi - lo > 1)
            {
                const unsigned mid = (lo + hi) >> 1;
                if (inclSmall[mid] <= limit) lo = mid;
                else hi = mid;
            }
            chunkEnd[d] = std::uint16_t(lo + 1);
        }
        __syncthreads();
        if (d == 0)
        {
            std::uint32_t n = 0;
            if (last)
            {
                chBegin[0] = 0;
                chLen[0] = par.size;
                n = 1;
            }
            else
                for (unsigned c = nextStart[0]; c < kRadix; c = nextStart[chunkEnd[c]])
                {
                    chBegin[n] = relS[c];
                    chLen[n] = relS[chunkEnd[c]] - relS[c];
                    ++n;
                }
            nChunks = n;
            const std::uint32_t nb = nBig, nrr = nRanges;
            if (root)
            {
                // the lists are empty: this parent's entries start at 0
                base[0] = base[1] = base[2] = 0;
                if (lead)
                {
                    cnt->segs[nextSet] = nb;
                    cnt->ranges[nextSet] = nrr;
                    cnt->chunks = n;
                }
            }
            else
            {
                // three independent atomics: one round trip
                const std::uint32_t b0 = nb ? atomicAdd(&cnt->segs[nextSet], nb) : 0u;
                const std::uint32_t b1 = nrr ? atomicAdd(&cnt->ranges[nextSet], nrr) : 0u;
                const std::uint32_t b2 = n ? atomicAdd(&cnt->chunks, n) : 0u;
                base[0] = b0;
                base[1] = b1;
                base[2] = b2;
            }
            if (lead && ((nb && base[0] + nb > maxSegs) || (nrr && base[1] + nrr > maxRanges) ||
                         (n && base[2] + n > maxChunks)))
                atomicAdd(&cnt->overflow, 1u);
        }
        __syncthreads();
        const bool fits = base[0] + nBig <= maxSegs && base[1] + nRanges <= maxRanges;
        if (lead && isBig && fits)
            nextSegs[base[0] + bigRank] = MsdSeg{par.begin + rel, s, base[1] + rangeRank, nr, 0};
        if (fits)
        {
            const std::uint32_t t0 = root ? blockIdx.x * kSortThreads : 0u;
            const std::uint32_t step = root ? gridDim.x * kSortThreads : kSortThreads;
            for (std::uint32_t t = t0 + d; t < nRanges; t += step)
            {
                // the big sub-bucket owning range t: last k with rangeRankOf[k] <= t
                std::uint32_t lo = 0, hi = nBig; // invariant: rangeRankOf[lo] <= t < rangeRankOf[hi]
                while (hi - lo > 1)
                {
                    const std::uint32_t mid = (lo + hi) >> 1;
                    if (rangeRankOf[mid] <= t) lo = mid;
                    else hi = mid;
                }
                const unsigned sb = bigIndex[lo];
                const std::uint32_t k = t - rangeRankOf[lo], o = k * kMsdTile;
                const std::uint32_t size = relS[sb + 1] - relS[sb];
                nextRanges[base[1] + t] = SegRange{par.begin + relS[sb] + o, size - o < kMsdTile ? size - o : kMsdTile, o,
                                                   base[0] + lo, 0};
                nextFirst[base[1] + t] = k == 0;
            }
        }
        if (lead && d < nChunks && base[2] + d < maxChunks)
            chunks[base[2] + d] = MsdChunk{par.begin + chBegin[d], chLen[d], bits, parity, 0};
        __syncthreads();
    }
}

// Digit histogram of the wave's byte range [a, e) into the wave's LDS table
// (16 digits per load in the aligned middle).
__device__ inline void countDigitBytes(const std::uint8_t *__restrict__ digits, std::size_t a, std::size_t e,
                                       std::uint32_t *h, unsigned lane)
{
    std::size_t a16 = (a + 15) & ~std::size_t(15);
    if (a16 > e) a16 = e;
    if (a + lane < a16) countDigit(h, digits[a + lane]);
    std::size_t e16 = e & ~std::size_t(15);
    if (e16 < a16) e16 = a16;
    const uint4 *quads = reinterpret_cast<const uint4 *>(digits);
    constexpr int kQuads = 4; // 16-byte loads in flight per lane (a 4096-key range in one step)
    for (std::size_t i0 = a16 + 16 * std::size_t(lane); i0 < e16; i0 += kQuads * 16 * 64)
    {
        uint4 q[kQuads];
#pragma unroll
        for (int u = 0; u < kQuads; ++u)
        {
            const std::size_t i = i0 + std::size_t(u) * 16 * 64;
            q[u] = i < e16 ? quads[i >> 4] : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < kQuads; ++u)
        {
            if (i0 + std::size_t(u) * 16 * 64 >= e16) break; // a lane stops at its first quad past the end
            const std::uint32_t w[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int b = 0; b < 4; ++b) countDigit(h, (w[k] >> (8 * b)) & 0xffu);
        }
    }
    if (e16 + lane < e) countDigit(h, digits[e16 + lane]);
}

// Digit histograms of the level's ranges, kUpRows per block (one wave each),
// into the digit-major table hist[d * stride + range].
__global__ __launch_bounds__(kUpRows * 64) void msdUpsweepKernel(const std::uint8_t *__restrict__ digits,
                                                                 const SegRange *__restrict__ ranges,
                                                                 const MsdCounters *cnt, int set,
                                                                 std::uint32_t *hist, std::size_t stride)
{
    __shared__ std::uint32_t h[kUpRows][kRadix];
    const std::uint32_t count = cnt->ranges[set] < stride ? cnt->ranges[set] : std::uint32_t(stride);
    const std::uint32_t groups = (count + kUpRows - 1) / kUpRows;
    const int wave = threadIdx.x >> 6;
    const unsigned lane = laneId();
    for (std::uint32_t g = blockIdx.x; g < groups; g += gridDim.x)
    {
        const std::size_t r0 = xcdRow(g, groups) * kUpRows;
        for (int i = threadIdx.x; i < kUpRows * kRadix; i += kUpRows * 64) (&h[0][0])[i] = 0;
        __syncthreads();
        const std::size_t r = r0 + std::size_t(wave);
        if (r < count)
        {
            const SegRange rg = ranges[r];
            countDigitBytes(digits, rg.begin, rg.begin + rg.len, h[wave], lane);
        }
        __syncthreads();
        if (threadIdx.x < kRadix)
        {
            const unsigned d = threadIdx.x;
            std::uint32_t *out = hist + std::size_t(d) * stride + r0;
            if (r0 + kUpRows <= count)
            {
                reinterpret_cast<uint4 *>(out)[0] = make_uint4(h[0][d], h[1][d], h[2][d], h[3][d]);
                reinterpret_cast<uint4 *>(out)[1] = make_uint4(h[4][d], h[5][d], h[6][d], h[7][d]);
            }
            else
                for (int w = 0; w < kUpRows; ++w)
                    if (r0 + w < count) out[w] = h[w][d];
        }
        __syncthreads();
    }
}

// Exclusive prefix of digit blockIdx.x's count over the ranges of each
// segment (a segmented scan; first[r] marks a segment's first range), in
// chunks of kScanChunk ranges: block (d, y) scans chunks y, y + gridDim.y,
// ... with 4 consecutive ranges per thread. A chunk that starts inside a
// segment first sums the digit's counts of that segment's earlier ranges
// as its carry. The count of the digit in a
// whole segment goes to subTotals[segment][digit]. Block (0, 0) also clears
// the list counters of the next level (its lists are no longer read).
constexpr int kScanPer = 4;
constexpr std::uint32_t kScanChunk = kScanPer * kSortThreads;
__global__ __launch_bounds__(kSortThreads) void msdScanKernel(const std::uint32_t *__restrict__ hist,
                                                              std::size_t stride,
                                                              const std::uint8_t *__restrict__ first,
                                                              const SegRange *__restrict__ ranges, MsdCounters *cnt,
                                                              int set, std::uint32_t *off, std::uint32_t *subTotals)
{
    constexpr int kPer = kScanPer;
    __shared__ std::uint32_t wS[kWaves];
    __shared__ std::uint32_t wF[kWaves];
    __shared__ std::uint32_t carryShared;
    const std::uint32_t count = cnt->ranges[set] < stride ? cnt->ranges[set] : std::uint32_t(stride);
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0)
    {
        cnt->segs[set ^ 1] = 0;
        cnt->ranges[set ^ 1] = 0;
    }
    const std::size_t d = blockIdx.x;
    const std::uint32_t *h = hist + d * stride;
    std::uint32_t *o = off + d * stride;
    const unsigned lane = laneId();
    const int wave = threadIdx.x >> 6;
    for (std::uint32_t r0 = blockIdx.y * kScanChunk; r0 < count; r0 += gridDim.y * kScanChunk)
    {
        // the chunk's own ranges (4-byte flag and 16-byte count loads: stride
        // and f0 are multiples of 4), loaded with the carry's inputs
        const std::uint32_t f0 = r0 + threadIdx.x * kPer;
        std::uint32_t v[kPer], ex[kPer];
        std::uint8_t fb[kPer];
        if (f0 + kPer <= count)
        {
            const std::uint32_t q = *reinterpret_cast<const std::uint32_t *>(first + f0);
            const uint4 c = *reinterpret_cast<const uint4 *>(h + f0);
            v[0] = c.x, v[1] = c.y, v[2] = c.z, v[3] = c.w;
#pragma unroll
            for (int k = 0; k < kPer; ++k) fb[k] = std::uint8_t(q >> (8 * k));
        }
        else
        {
#pragma unroll
            for (int k = 0; k < kPer; ++k)
            {
                const bool in = f0 + k < count;
                v[k] = in ? h[f0 + k] : 0u;
                fb[k] = in ? first[f0 + k] : 0;
            }
        }
        const bool nextFirst = f0 + kPer >= count || first[f0 + kPer];
        // carry: the digit's count in the ranges of the segment open at r0
        // that precede the chunk (block-uniform branch); range k of a
        // segment starts kMsdTile * k keys into it
        std::uint32_t part = 0;
        const std::uint32_t into = ranges[r0].localBase / kMsdTile;
        for (std::uint32_t r = r0 - into + threadIdx.x; r < r0; r += kSortThreads) part += h[r];
#pragma unroll
        for (int o2 = 32; o2 >= 1; o2 >>= 1) part += __shfl_xor(part, o2, 64);
        std::uint32_t flags = 0; // bit k: range f0+k starts a segment
        std::uint32_t run = 0;
#pragma unroll
        for (int k = 0; k < kPer; ++k)
        {
            const bool in = f0 + k < count;
            const bool f = in && fb[k];
            flags |= std::uint32_t(f) << k;
            if (f) run = 0;
            ex[k] = run;
            run += v[k];
        }
        // inclusive segmented scan of the thread aggregates (run, has flag) in the wave
        std::uint32_t s = run, fl = flags != 0;
#pragma unroll
        for (int sh = 1; sh < 64; sh <<= 1)
        {
            const std::uint32_t ps = __shfl_up(s, sh, 64), pf = __shfl_up(fl, sh, 64);
            if (lane >= unsigned(sh))
            {
                if (!fl) s += ps;
                fl |= pf;
            }
        }
        std::uint32_t es = __shfl_up(s, 1, 64), ef = __shfl_up(fl, 1, 64); // exclusive within the wave
        if (lane == 0) es = 0, ef = 0;
        if (lane == 63)
        {
            wS[wave] = s;
            wF[wave] = fl;
        }
        if (threadIdx.x == 0) carryShared = 0;
        __syncthreads();
        if (lane == 0 && part) atomicAdd(&carryShared, part);
        __syncthreads();
        const std::uint32_t carry = carryShared;
        // carry into this wave: the chunk's carry, then the earlier waves' totals
        std::uint32_t cs = carry;
        for (int w = 0; w < wave; ++w)
        {
            const std::uint32_t ws = wS[w], wf = wF[w];
            cs = wf ? ws : cs + ws;
        }
        const std::uint32_t tin = ef ? es : cs + es;
        bool seen = false;
#pragma unroll
        for (int k = 0; k < kPer; ++k)
        {
            if ((flags >> k) & 1u) seen = true;
            const std::uint32_t val = seen ? ex[k] : tin + ex[k];
            const std::uint32_t r = f0 + k;
            if (r < count)
            {
                o[r] = val;
                const bool end = k + 1 < kPer ? ((flags >> (k + 1)) & 1u) != 0 || r + 1 == count : nextFirst;
                if (end) subTotals[std::size_t(ranges[r].seg) * kRadix + d] = val + v[k];
            }
        }
        __syncthreads(); // wS, wF and carryShared are rewritten by the next chunk
    }
}

// A value that is the same on every lane, kept in scalar registers.
__device__ inline Key uniformKey(Key v)
{
    const unsigned lo = unsigned(__builtin_amdgcn_readfirstlane(int(unsigned(v))));
    const unsigned hi = unsigned(__builtin_amdgcn_readfirstlane(int(unsigned(v >> 32))));
    return Key(lo) | (Key(hi) << 32);
}

// Sorts every chunk in LDS. The chunk's keys (and values) are held in
// registers; a stable counting pass per 8-bit digit (ranking as in
// scatterKernel) works on d = key - base, base the chunk's smallest key (d
// orders the chunk's keys as the keys do), and covers the kFixBits highest
// bits of the chunk's span: from the span's highest bit `top` down to
// fix = top + 1 - kFixBits (two windows of 8 bits, which need not be byte
// aligned). A chunk of 2048 keys spread evenly over its span then has runs
// of 2048 / 2^kFixBits keys on average that agree above `fix`, whatever its
// depth in the key space; the runs are finished on the lower bits: pairs by
// one compare-and-swap, runs of up to kFixRun keys by an insertion sort (one
// thread per run), runs of up to kWaveRun keys by a rank sort of one wave
// per run. If a run is longer, the chunk takes a counting pass on the 8 bits
// below `fix`, the high windows again (LSD order) and the same fix-up from
// bit fix - 8; if that fails too, every bit of the span gets a counting pass
// (plain LSD order). Chunks with bits == 0 are already in order and are only
// copied when they lie in the other buffer.
constexpr std::uint32_t kFixRun = 16;
constexpr std::uint32_t kWaveRun = 64;
constexpr int kFixBits = 2 * kRadixBits; // bits sorted by counting passes before the fix-up
constexpr std::uint32_t kRunCap = kMsdChunk / 3 + 1; // runs of 3 or more keys in a chunk

__global__ __launch_bounds__(kLocThreads, 4) void msdLocalSortKernel(const MsdCounters *cnt,
                                                                     const MsdChunk *__restrict__ chunks, Key *X,
                                                                     Key *Y, std::uint32_t *vX, std::uint32_t *vY,
                                                                     int finalIsY, std::uint32_t maxChunks,
                                                                     const std::uint32_t *fromChunk)
{
    constexpr int kW = kLocThreads / 64;
    __shared__ Key stageKeys[kMsdChunk];
    __shared__ std::uint32_t stageVals[kMsdChunk];
    __shared__ std::uint32_t waveCount[kW][kRadix];
    __shared__ std::uint32_t sink[kW][64];
    __shared__ std::uint32_t scratch[kRadix / 64];
    __shared__ Key waveLo[kW], waveHi[kW];
    __shared__ std::uint32_t longRun;   // a run longer than kWaveRun was seen
    __shared__ std::uint32_t shortRuns; // runs of 3 .. kFixRun keys, from the front of runList
    __shared__ std::uint32_t waveRuns;  // runs of kFixRun+1 .. kWaveRun keys, from the back
    __shared__ std::uint32_t runList[kRunCap]; // (start << 8) | length
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const unsigned lane = laneId();
    Key *kd = finalIsY ? Y : X;
    std::uint32_t *vd = finalIsY ? vY : vX;
    const std::uint32_t count = cnt->chunks < maxChunks ? cnt->chunks : maxChunks;
    const std::uint32_t from = fromChunk ? *fromChunk : 0u; // chunks before it were sorted by an earlier launch
    const std::uint32_t first = std::uint32_t(wave) * (kLocItems * 64) + lane;
    Key key[kLocItems];
    std::uint32_t val[kLocItems];
    Key base = 0; // the chunk's smallest key; the registers and the stage hold key - base

    // one stable counting pass on the digit of key - base at `shift`; the
    // result is in the registers (striped order) and in stageKeys / stageVals
    auto pass = [&](int shift) __attribute__((always_inline)) {
        for (int i = tid; i < kW * kRadix; i += kLocThreads) (&waveCount[0][0])[i] = 0;
        __syncthreads();
        std::uint32_t rank[kLocItems], old[kLocItems];
        int leader[kLocItems];
#pragma unroll
        for (int j = 0; j < kLocItems; ++j)
        {
            const unsigned dg = unsigned(key[j] >> shift) & (kRadix - 1);
            const std::uint64_t match = laneMatch(dg);
            const std::uint32_t lower = lanesBelow(match);
            const bool lead = lower == 0;
            std::uint32_t *slot = lead ? &waveCount[wave][dg] : &sink[wave][lane];
            old[j] = atomicAdd(slot, lead ? std::uint32_t(__popcll(match)) : 0u);
            leader[j] = __builtin_ctzll(match);
            rank[j] = lower;
        }
#pragma unroll
        for (int j = 0; j < kLocItems; ++j) rank[j] += std::uint32_t(__shfl(int(old[j]), leader[j], 64));
        __syncthreads();
        std::uint32_t tileCount = 0, inc = 0;
        if (tid < kRadix)
        {
#pragma unroll
            for (int w = 0; w < kW; ++w)
            {
                const std::uint32_t cw = waveCount[w][tid];
                waveCount[w][tid] = tileCount;
                tileCount += cw;
            }
            inc = tileCount;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1)
            {
                const std::uint32_t t = __shfl_up(inc, o, 64);
                if (lane >= unsigned(o)) inc += t;
            }
            if (lane == 63) scratch[wave] = inc;
        }
        __syncthreads();
        if (tid < kRadix)
        {
            std::uint32_t before = 0;
#pragma unroll
            for (int w = 0; w < kRadix / 64; ++w)
                if (w < wave) before += scratch[w];
            const std::uint32_t start = before + inc - tileCount;
#pragma unroll
            for (int w = 0; w < kW; ++w) waveCount[w][tid] += start;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kLocItems; ++j)
        {
            const std::uint32_t pos = rank[j] + waveCount[wave][unsigned(key[j] >> shift) & (kRadix - 1)];
            stageKeys[pos] = key[j];
            stageVals[pos] = val[j];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kLocItems; ++j)
        {
            key[j] = stageKeys[first + std::uint32_t(j) * 64];
            val[j] = stageVals[first + std::uint32_t(j) * 64];
        }
        __syncthreads();
    };

    for (std::uint32_t c = from + blockIdx.x; c < count; c += gridDim.x)
    {
        const MsdChunk ch = chunks[c];
        const Key *ks = ch.parity ? Y : X;
        const std::uint32_t *vs = ch.parity ? vY : vX;
        if (ch.bits == 0 || ch.len > kMsdChunk || ch.len < 2)
        {
            if (ks != kd)
                for (std::uint32_t i = tid; i < ch.len; i += kLocThreads)
                {
                    kd[ch.begin + i] = ks[ch.begin + i];
                    vd[ch.begin + i] = vs[ch.begin + i];
                }
            continue;
        }
        Key lo = ~Key(0), hi = 0;
#pragma unroll
        for (int j = 0; j < kLocItems; ++j)
        {
            const std::uint32_t idx = first + std::uint32_t(j) * 64;
            const bool in = idx < ch.len;
            const std::uint32_t cidx = in ? idx : 0u;
            const Key k = ks[ch.begin + cidx];
            key[j] = k;
            val[j] = vs[ch.begin + cidx];
            if (in)
            {
                lo = k < lo ? k : lo;
                hi = k > hi ? k : hi;
            }
        }
        // the chunk's smallest and largest keys
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
        {
            const Key a = Key(__shfl_xor((unsigned long long)lo, o, 64));
            const Key b = Key(__shfl_xor((unsigned long long)hi, o, 64));
            lo = a < lo ? a : lo;
            hi = b > hi ? b : hi;
        }
        if (lane == 0) waveLo[wave] = lo, waveHi[wave] = hi;
        __syncthreads();
        lo = waveLo[0], hi = waveHi[0];
#pragma unroll
        for (int w = 1; w < kW; ++w)
        {
            lo = waveLo[w] < lo ? waveLo[w] : lo;
            hi = waveHi[w] > hi ? waveHi[w] : hi;
        }
        base = lo;
        const Key span = hi - lo; // block-uniform
        const int top = span ? 63 - __builtin_clzll(span) : 0;
        const Key dmask = top == 63 ? ~Key(0) : (Key(2) << top) - 1; // the bits key - base may use
        // from here on the registers hold key - base; padding: dmask, the
        // largest value in every digit (it still ranks last)
#pragma unroll
        for (int j = 0; j < kLocItems; ++j) key[j] = first + std::uint32_t(j) * 64 < ch.len ? key[j] - base : dmask;
        // Fix-up of the runs of keys whose key - base agree from bit `shift`
        // up, with the chunk sorted on those bits in the stage. Returns false,
        // changing nothing, when a run is longer than kWaveRun.
        auto fixRuns = [&](int shift) __attribute__((always_inline)) -> bool {
            if (tid == 0) longRun = 0, shortRuns = 0, waveRuns = 0;
            __syncthreads();
            auto high = [&](std::uint32_t q) { return stageKeys[q] >> shift; };
            // position p = t + k*kLocThreads of thread t: a wave's lanes hold
            // 64 consecutive positions, so a key's successor is the next
            // lane's (the row's last lane reads the next row's first key), and
            // a run's length inside the row follows from one ballot; only a
            // run that leaves the row walks on in LDS. Nothing is written
            // before the barrier: every key's high bits stay as read.
            std::uint32_t pairs = 0; // bit k: position t + k*kLocThreads starts a run of 2
#pragma unroll 4
            for (int k = 0; k < kLocItems; ++k)
            {
                const std::uint32_t p = std::uint32_t(tid) + std::uint32_t(k) * kLocThreads;
                const std::uint32_t row = p - lane; // the row's first position (wave-uniform)
                const Key kp = p < ch.len ? stageKeys[p] : Key(0);
                const Key nextRow = uniformKey(row + 64 < ch.len ? stageKeys[row + 64] : Key(0));
                const Key prevRow = uniformKey(row > 0 ? stageKeys[row - 1] : Key(0));
                Key kn = Key(__shfl_down((unsigned long long)kp, 1, 64));
                if (lane == 63) kn = nextRow;
                const Key h = kp >> shift;
                const bool eqNext = p + 1 < ch.len && (kn >> shift) == h;
                const std::uint64_t nextMask = __ballot(eqNext);
                const bool eqPrev = lane > 0 ? ((nextMask >> (lane - 1)) & 1u) != 0
                                             : p > 0 && p < ch.len && (prevRow >> shift) == h;
                if (eqNext && !eqPrev)
                {
                    // bits lane.. of nextMask: the run's positions that equal their successor
                    const std::uint64_t rest = ~(nextMask >> lane);
                    const std::uint32_t ones = rest ? std::uint32_t(__builtin_ctzll(rest)) : 64u;
                    std::uint32_t e = p + ones + 1;
                    if (ones == 64u - lane)
                        while (e < ch.len && e - p <= kWaveRun && high(e) == h) ++e;
                    const std::uint32_t len = e - p;
                    if (len == 2) pairs |= 1u << k;
                    else if (len > kWaveRun) longRun = 1;
                    else if (len > kFixRun) runList[kRunCap - 1 - atomicAdd(&waveRuns, 1u)] = (p << 8) | len;
                    else runList[atomicAdd(&shortRuns, 1u)] = (p << 8) | len;
                }
            }
            __syncthreads();
            if (longRun) return false; // block-uniform
            // pairs: one compare-and-swap (stable: equal keys stay)
#pragma unroll 1
            for (std::uint32_t bitsLeft = pairs; bitsLeft; bitsLeft &= bitsLeft - 1)
                {
                    const std::uint32_t p = std::uint32_t(tid) + std::uint32_t(__builtin_ctz(bitsLeft)) * kLocThreads;
                    const Key a = stageKeys[p], b = stageKeys[p + 1];
                    if (b < a)
                    {
                        const std::uint32_t va = stageVals[p], vb = stageVals[p + 1];
                        stageKeys[p] = b;
                        stageKeys[p + 1] = a;
                        stageVals[p] = vb;
                        stageVals[p + 1] = va;
                    }
                }
            // runs of 3..kFixRun keys: a stable insertion sort, one thread per
            // run (keys of a run agree above `shift`, so whole keys compare
            // as their low bits)
            const std::uint32_t ns = shortRuns;
#pragma unroll 1
            for (std::uint32_t r = std::uint32_t(tid); r < ns; r += kLocThreads)
            {
                const std::uint32_t e = runList[r], p0 = e >> 8, end = p0 + (e & 0xffu);
                for (std::uint32_t i = p0 + 1; i < end; ++i)
                {
                    const Key kk = stageKeys[i];
                    const std::uint32_t vv = stageVals[i];
                    std::uint32_t q = i;
                    while (q > p0 && stageKeys[q - 1] > kk)
                    {
                        stageKeys[q] = stageKeys[q - 1];
                        stageVals[q] = stageVals[q - 1];
                        --q;
                    }
                    stageKeys[q] = kk;
                    stageVals[q] = vv;
                }
            }
            // longer runs: one wave each; lane i ranks the run's key i against
            // all keys of the run (ties by position: stable)
            const std::uint32_t nw = waveRuns;
#pragma unroll 1
            for (std::uint32_t r = std::uint32_t(wave); r < nw; r += kW)
            {
                const std::uint32_t e = runList[kRunCap - 1 - r], p0 = e >> 8, len = e & 0xffu;
                const bool in = lane < len;
                const Key kk = in ? stageKeys[p0 + lane] : Key(0);
                const std::uint32_t vv = in ? stageVals[p0 + lane] : 0u;
                std::uint32_t rank = 0;
#pragma unroll 1
                for (std::uint32_t j = 0; j < len; ++j)
                {
                    // readlane returns int: widen through unsigned (no sign extension)
                    const Key kj = Key(unsigned(__builtin_amdgcn_readlane(int(unsigned(kk)), int(j)))) |
                                   (Key(unsigned(__builtin_amdgcn_readlane(int(unsigned(kk >> 32)), int(j)))) << 32);
                    rank += (kj < kk || (kj == kk && j < lane)) ? 1u : 0u;
                }
                if (in)
                {
                    stageKeys[p0 + rank] = kk;
                    stageVals[p0 + rank] = vv;
                }
            }
            __syncthreads();
            return true;
        };
        auto fromStage = [&]() __attribute__((always_inline)) {
#pragma unroll
            for (int j = 0; j < kLocItems; ++j)
            {
                key[j] = stageKeys[first + std::uint32_t(j) * 64];
                val[j] = stageVals[first + std::uint32_t(j) * 64];
            }
        };
        // counting passes on the digits of key - base in [lo, 64) that mask
        // covers, lowest first
        auto passesFrom = [&](int lo, Key mask) __attribute__((always_inline)) {
            for (int shift = lo; shift < 64; shift += kRadixBits)
                if ((mask >> shift) & (kRadix - 1)) pass(shift);
        };
        if (span)
        {
            // passes on [fix, top], the kFixBits highest bits of the span;
            // they leave the result in the registers and the stage
            const int fix = top >= kFixBits ? top + 1 - kFixBits : 0;
            passesFrom(fix, dmask & ~((Key(1) << fix) - 1));
            if (fix > 0)
            {
                bool done = fixRuns(fix);
                if (!done)
                {
                    // A run longer than kWaveRun (duplicate-heavy keys). The
                    // fallback passes take only the bits of key - base that
                    // vary within the chunk -- their OR, as the smallest key's
                    // is 0 -- skipping windows in which no key differs (a
                    // chunk of two distinct values takes one pass), with the
                    // padding set to that OR (the largest value in every
                    // digit, equal to the keys on the other bits).
                    Key vary = 0;
#pragma unroll
                    for (int j = 0; j < kLocItems; ++j)
                        if (first + std::uint32_t(j) * 64 < ch.len) vary |= key[j];
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) vary |= Key(__shfl_xor((unsigned long long)vary, o, 64));
                    if (lane == 0) waveLo[wave] = vary;
                    __syncthreads(); // also: the passes below rewrite the stage
                    vary = 0;
#pragma unroll
                    for (int w = 0; w < kW; ++w) vary |= waveLo[w];
#pragma unroll
                    for (int j = 0; j < kLocItems; ++j)
                        if (first + std::uint32_t(j) * 64 >= ch.len) key[j] = vary;
                    // the 8 bits below fix, then the high windows again
                    const int fix2 = fix > kRadixBits ? fix - kRadixBits : 0;
                    const Key below2 = (Key(1) << fix2) - 1;
                    passesFrom(fix2, vary & ~below2);
                    done = (vary & below2) == 0 || fixRuns(fix2);
                    if (!done)
                    {
                        __syncthreads();
                        passesFrom(0, vary);
                    }
                }
                if (done) fromStage();
            }
        }
#pragma unroll
        for (int j = 0; j < kLocItems; ++j)
        {
            const std::uint32_t idx = first + std::uint32_t(j) * 64;
            if (idx < ch.len)
            {
                kd[ch.begin + idx] = key[j] + base;
                vd[ch.begin + idx] = val[j];
            }
        }
        __syncthreads(); // waveLo / waveHi, the run lists and the stage are rewritten by the next chunk
    }
}

template <int THREADS, int ITEMS, bool PRE = false>
void launchScatter(hipStream_t s, const SortPlan &pl, SortWorkspace &ws, const Key *kin, Key *kout,
                   const std::uint32_t *vin, std::uint32_t *vout, bool vals, bool iota, bool digits, int shift,
                   int next)
{
    const dim3 grid(unsigned(pl.ranges)), block(THREADS);
#define AT_SCATTER(V, I, D)                                                                                          \
    hipLaunchKernelGGL((scatterKernel<THREADS, ITEMS, V, I, D, false, PRE>), grid, block, 0, s, kin, kout, vin, vout, \
                       ws.digits, pl.n, pl.rangeKeys, ws.rangeOff, ws.digitTotal, pl.ranges, shift, next)
    if (!vals)
    {
        if (digits) AT_SCATTER(false, false, true);
        else AT_SCATTER(false, false, false);
    }
    else if (iota)
    {
        if (digits) AT_SCATTER(true, true, true);
        else AT_SCATTER(true, true, false);
    }
    else
    {
        if (digits) AT_SCATTER(true, false, true);
        else AT_SCATTER(true, false, false);
    }
#undef AT_SCATTER
}

void scanOffsets(hipStream_t s, const SortPlan &pl, SortWorkspace &ws)
{
    hipLaunchKernelGGL(scanDigitKernel, dim3(kRadix), dim3(kSortThreads), 0, s, ws.rangeHist, pl.ranges, pl.parts,
                       ws.rangeOff, ws.digitTotal);
}

// Passes 0..7 given pass 0's range histograms in ws.rangeHist. Keys ping-pong
// between kA (input, and output after the even number of passes) and kB.
// One scatter pass of the plan's tile shape.
void dispatchScatter(hipStream_t s, const SortPlan &pl, SortWorkspace &ws, const Key *src, Key *dst,
                     const std::uint32_t *vsrc, std::uint32_t *vdst, bool vals, bool first, bool digits, int shift,
                     int next)
{
    // next-tile prefetch is instantiated for tiles of 8 or 12 keys per thread
    const bool pre = g_prefetch && pl.items <= 12 && pl.threads <= 512;
#define AT_ARGS s, pl, ws, src, dst, vsrc, vdst, vals, first, digits, shift, next
    switch (pl.threads * 100 + pl.items + (pre ? 1000000 : 0))
    {
    case 1025608: launchScatter<256, 8, true>(AT_ARGS); break;
    case 1025612: launchScatter<256, 12, true>(AT_ARGS); break;
    case 1051208: launchScatter<512, 8, true>(AT_ARGS); break;
    case 1051212: launchScatter<512, 12, true>(AT_ARGS); break;
    case 25608: launchScatter<256, 8>(AT_ARGS); break;
    case 25612: launchScatter<256, 12>(AT_ARGS); break;
    case 25616: launchScatter<256, 16>(AT_ARGS); break;
    case 51208: launchScatter<512, 8>(AT_ARGS); break;
    case 51212: launchScatter<512, 12>(AT_ARGS); break;
    case 51216: launchScatter<512, 16>(AT_ARGS); break;
    case 102408: launchScatter<1024, 8>(AT_ARGS); break;
    default: throw Error(1, "no scatter kernel for this tile shape");
    }
#undef AT_ARGS
}

void runPasses(hipStream_t s, const SortPlan &pl, SortWorkspace &ws, Key *kA, Key *kB, std::uint32_t *vA,
               std::uint32_t *vB, bool vals, bool iota, bool firstHistWhole, bool timed = true)
{
    for (int pass = 0; pass < kSortPasses; ++pass)
    {
        const Key *src = (pass & 1) ? kB : kA;
        Key *dst = (pass & 1) ? kA : kB;
        const std::uint32_t *vsrc = (pass & 1) ? vB : vA;
        std::uint32_t *vdst = (pass & 1) ? vA : vB;
        const int shift = pass * kRadixBits;
        if (pass > 0)
        {
            prof::Scope scope("sort_upsweep", s, timed);
            if (g_digitArray)
                hipLaunchKernelGGL(upsweepDigitsKernel, dim3(unsigned((pl.ranges + kUpRows - 1) / kUpRows)),
                                   dim3(kUpRows * 64), 0, s, ws.digits, pl.n, pl.rangeKeys, pl.ranges, ws.rangeHist);
            else
                hipLaunchKernelGGL(upsweepKeysKernel, dim3(unsigned(pl.ranges)), dim3(kSortThreads), 0, s, src,
                                   pl.n, pl.rangeKeys, shift, ws.rangeHist);
        }
        // key generation writes `parts` partial histograms per range;
        // every other histogram has one row per range
        SortPlan scanPlan = pl;
        if (pass > 0 || firstHistWhole) scanPlan.parts = 1;
        scanOffsets(s, scanPlan, ws);
        const bool digits = g_digitArray && pass + 1 < kSortPasses;
        prof::Scope scope("sort_pass", s, timed);
        const bool first = iota && pass == 0;
        dispatchScatter(s, pl, ws, src, dst, vsrc, vdst, vals, first, digits, shift, shift + kRadixBits);
    }
    AT_HIP(hipGetLastError());
}

// The 11-bit LSD passes when they are chosen and the workspace holds their
// histograms (fewer than kWideMaxKeys keys).
bool widePasses(const SortWorkspace &ws, std::size_t n)
{
    return g_digitBits == kWideBits && ws.wideHist && (n + kWideTile - 1) / kWideTile <= ws.wideRanges;
}
void runWidePasses(hipStream_t s, std::size_t n, SortWorkspace &ws, Key *kA, Key *kB, std::uint32_t *vA,
                   std::uint32_t *vB, bool vals, bool iota)
{
    const std::uint32_t ranges = std::uint32_t((n + kWideTile - 1) / kWideTile);
    for (int pass = 0; pass < kWidePasses; ++pass)
    {
        const Key *src = (pass & 1) ? kB : kA;
        Key *dst = (pass & 1) ? kA : kB;
        const std::uint32_t *vsrc = (pass & 1) ? vB : vA;
        std::uint32_t *vdst = (pass & 1) ? vA : vB;
        const int shift = pass * kWideBits;
        {
            prof::Scope scope("sort_upsweep", s);
            hipLaunchKernelGGL(wideUpsweepKernel, dim3(ranges), dim3(kWideThreads), 0, s, src, n, shift, ws.wideHist,
                               ranges);
            hipLaunchKernelGGL(wideScanKernel, dim3(kWideRadix), dim3(kWideThreads), 0, s, ws.wideHist, ws.wideOff,
                               ws.wideTotal, ranges);
        }
        prof::Scope scope("sort_pass", s);
        const dim3 grid(ranges), block(kWideThreads);
        if (!vals)
            hipLaunchKernelGGL((wideScatterKernel<false, false>), grid, block, 0, s, src, dst, nullptr, nullptr, n,
                               shift, ws.wideOff, ws.wideTotal, ranges);
        else if (iota && pass == 0)
            hipLaunchKernelGGL((wideScatterKernel<true, true>), grid, block, 0, s, src, dst, nullptr, vdst, n, shift,
                               ws.wideOff, ws.wideTotal, ranges);
        else
            hipLaunchKernelGGL((wideScatterKernel<true, false>), grid, block, 0, s, src, dst, vsrc, vdst, n, shift,
                               ws.wideOff, ws.wideTotal, ranges);
    }
    AT_HIP(hipGetLastError());
}

// Copies a device count into (pinned, coherent) host memory for a polling host.
constexpr std::uint32_t kCountPending = 0xffffffffu;
__global__ void publishCountKernel(const std::uint32_t *count, std::uint32_t *host)
{
    __atomic_store_n(host, *count, __ATOMIC_RELEASE);
    __threadfence_system();
}
// Records the chunk count (the chunks an early chunk-sort launch takes).
__global__ void markChunksKernel(MsdCounters *c) { c->pad[0] = c->chunks; }

// Levels 2..8 and the LDS chunk sort, after level 1 has partitioned the
// keys (X -> Y, next-digit bytes in `digits`) and planned level 2's lists
// (set 0) and level 1's chunks. Odd levels write Y, even levels X; the result
// ends in Y when finalIsY, else in X.
// hostPoll = false keeps the fixed launch sequence outside capture too (the
// device-driven batch build, whose host waits once, after its last pass).
void msdLevels(hipStream_t s, SortWorkspace &ws, std::size_t n, std::uint8_t *digits, Key *X, Key *Y,
               std::uint32_t *vX, std::uint32_t *vY, bool finalIsY, bool hostPoll = true)
{
    const std::uint32_t maxSegs = std::uint32_t(ws.msdMaxSegs), maxRanges = std::uint32_t(ws.msdStride),
                        maxChunks = std::uint32_t(ws.msdMaxChunks);
    // Outside graph capture the host may look at the number of segments of
    // level 4 (one small read): when it is zero, levels 4..8 have no work and
    // are not launched. A captured sort keeps the fixed launch sequence.
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    AT_HIP(hipStreamIsCapturing(s, &capture));
    // about one block per chunk (spread keys give fewer than 2 per
    // kMsdChunk keys; the loop covers any excess): blocks retire and start
    // independently, so their load, sort and store phases overlap
    const unsigned localGrid = unsigned(2 * (n / kMsdChunk) + 256);
    auto localSort = [&](const std::uint32_t *fromChunk) {
        prof::Scope scope("msd_local", s);
        hipLaunchKernelGGL(msdLocalSortKernel, dim3(localGrid), dim3(kLocThreads), 0, s, ws.msdCnt, ws.msdChunks, X, Y,
                           vX, vY, int(finalIsY), maxChunks, fromChunk);
    };
    const std::uint32_t *from = nullptr;
    for (int level = 2; level <= kSortPasses; ++level)
    {
        if (level == 4 && hostPoll && capture == hipStreamCaptureStatusNone)
        {
            // Levels 4..8 have work only for keys that share 24 high bits in
            // buckets of more than kMsdChunk. Outside graph capture the
            // chunks planned so far are sorted at once, and the host reads
            // level 4's segment count meanwhile (a one-thread kernel stores
            // it into pinned host memory, polled; a stream synchronisation
            // after a bound): when it is zero the sort is complete, otherwise
            // the remaining levels and a chunk-sort launch for their chunks
            // follow. A captured sort keeps the fixed launch sequence.
            volatile std::uint32_t *flag = ws.msdHostCount;
            *flag = kCountPending;
            hipLaunchKernelGGL(publishCountKernel, dim3(1), dim3(1), 0, s, &ws.msdCnt->segs[level & 1],
                               ws.msdHostCount);
            hipLaunchKernelGGL(markChunksKernel, dim3(1), dim3(1), 0, s, ws.msdCnt);
            localSort(nullptr);
            AT_HIP(hipGetLastError());
            const auto t0 = std::chrono::steady_clock::now();
            while (*flag == kCountPending &&
                   std::chrono::steady_clock::now() - t0 < std::chrono::milliseconds(20))
            {
            }
            if (*flag == kCountPending) syncStream(s);
            if (*flag == kCountPending) throw Error(4, "sort: a level count never arrived");
            if (*flag == 0) return;
            from = &ws.msdCnt->pad[0];
        }
        prof::Scope scope("msd_level", s);
        // levels past the third see work only for keys that share 24 high
        // bits in buckets of more than kMsdChunk (heavy duplicates): small
        // grids keep those launches cheap when they are empty
        const unsigned cap = level <= 3 ? 1024u : 128u;
        const unsigned planCap = level <= 3 ? 2048u : 128u; // one parent per block (8 blocks per CU fit)
        const unsigned segGrid = unsigned(maxSegs < planCap ? maxSegs : planCap);
        const unsigned upGrid = unsigned((maxRanges / kUpRows) < cap ? maxRanges / kUpRows + 1 : cap);
        // levels 2 and 3 (the ones with work on spread keys): about one
        // block per range (at most n / tile + 256^(level - 1) ranges; the
        // loop covers any excess), later levels a small looping grid
        const std::size_t tiles = n / kMsdTile + 1;
        const unsigned scatterGrid =
            level == 2 ? unsigned(tiles + 256) : level == 3 ? unsigned(tiles + 4096) : 256u;
        const int set = level & 1;
        const bool srcY = (level - 1) & 1, dstY = level & 1;
        const Key *src = srcY ? Y : X;
        Key *dst = dstY ? Y : X;
        const std::uint32_t *vsrc = srcY ? vY : vX;
        std::uint32_t *vdst = dstY ? vY : vX;
        const int shift = 64 - kRadixBits * level;
        hipLaunchKernelGGL(msdUpsweepKernel, dim3(upGrid), dim3(kUpRows * 64), 0, s, digits, ws.msdRanges[set],
                           ws.msdCnt, set, ws.msdHist, ws.msdStride);
        const unsigned scanChunks = unsigned((ws.msdStride + kScanChunk - 1) / kScanChunk);
        hipLaunchKernelGGL(msdScanKernel, dim3(kRadix, scanChunks < kMsdScanRows ? scanChunks : kMsdScanRows),
                           dim3(kSortThreads), 0, s, ws.msdHist, ws.msdStride, ws.msdFirst[set], ws.msdRanges[set],
                           ws.msdCnt, set, ws.msdOff, ws.msdSubTotals);
        const MsdScatterArgs a{&ws.msdCnt->ranges[set], ws.msdSegs[set], ws.msdSubTotals, ws.msdOff, ws.msdStride};
        if (level < kSortPasses)
            hipLaunchKernelGGL((scatterKernel<kMsdThreads, kMsdItems, true, false, true, true, false, true>),
                               dim3(scatterGrid), dim3(kMsdThreads), 0, s, src, dst, vsrc, vdst, digits, n,
                               std::size_t(0), nullptr, nullptr, std::size_t(0), shift, shift - kRadixBits,
                               ws.msdRanges[set], a);
        else
            hipLaunchKernelGGL((scatterKernel<kMsdThreads, kMsdItems, true, false, false, true, false, true>),
                               dim3(scatterGrid), dim3(kMsdThreads), 0, s, src, dst, vsrc, vdst, digits, n,
                               std::size_t(0), nullptr, nullptr, std::size_t(0), shift, 0, ws.msdRanges[set], a);
        hipLaunchKernelGGL(msdPlanKernel, dim3(segGrid), dim3(kSortThreads), 0, s, level, ws.msdSegs[set],
                           &ws.msdCnt->segs[set], ws.msdSubTotals, ws.msdCnt, set ^ 1, ws.msdSegs[set ^ 1],
                           ws.msdRanges[set ^ 1], ws.msdFirst[set ^ 1], maxSegs, maxRanges, ws.msdChunks,
                           maxChunks, maxSegs);
    }
    localSort(from);
    AT_HIP(hipGetLastError());
}

// MSD sort of (key, value) pairs given level 1's range histograms of bits
// 56..63 in ws.rangeHist. Keys start in X (values in vX, or slots when
// iota); odd levels write Y, even levels X; the result ends in Y when
// finalIsY, else in X.
void msdSort(hipStream_t s, const SortPlan &pl, SortWorkspace &ws, Key *X, Key *Y, std::uint32_t *vX,
             std::uint32_t *vY, bool iota, bool firstHistWhole, bool finalIsY)
{
    const std::size_t n = pl.n;
    ws.allocateMsd(n, 0, s);
    hipLaunchKernelGGL(msdResetKernel, dim3(1), dim3(1), 0, s, ws.msdCnt, ws.msdRoot, n);
    SortPlan scanPlan = pl;
    if (firstHistWhole) scanPlan.parts = 1;
    scanOffsets(s, scanPlan, ws);
    {
        prof::Scope scope("sort_pass", s);
        dispatchScatter(s, pl, ws, X, Y, vX, vY, true, iota, true, 56, 48);
    }
    hipLaunchKernelGGL(msdPlanKernel, dim3(kRootPlanBlocks), dim3(kSortThreads), 0, s, 1, ws.msdRoot, nullptr,
                       ws.digitTotal,
                       ws.msdCnt, 0, ws.msdSegs[0], ws.msdRanges[0], ws.msdFirst[0], std::uint32_t(ws.msdMaxSegs),
                       std::uint32_t(ws.msdStride), ws.msdChunks, std::uint32_t(ws.msdMaxChunks), 1u);
    msdLevels(s, ws, n, ws.digits, X, Y, vX, vY, finalIsY);
}
} // namespace

int sortItemsPerThread() { return g_items; }
void setSortItemsPerThread(int items)
{
    if (items != 8 && items != 12 && items != 16) throw Error(1, "sort items per thread must be 8, 12 or 16");
    g_items = items;
}
int sortThreads() { return g_threads; }
void setSortThreads(int threads)
{
    if (threads != 256 && threads != 512 && threads != 1024) throw Error(1, "sort threads must be 256, 512 or 1024");
    g_threads = threads;
}
int sortPrefetch() { return g_prefetch; }
void setSortPrefetch(int on)
{
    if (on != 0 && on != 1) throw Error(1, "sort_prefetch must be 0 or 1");
    g_prefetch = on;
}
int sortMsd() { return g_msd; }
void setSortMsd(int on)
{
    if (on != 0 && on != 1) throw Error(1, "sort_msd must be 0 or 1");
    g_msd = on;
}
std::size_t sortMsdMinKeys() { return g_msdMinKeys; }
void setSortMsdMinKeys(std::size_t n)
{
    if (n < kMsdChunk || n >= (std::size_t(1) << 31))
        throw Error(1, "sort_msd_min_keys must be in [" + std::to_string(kMsdChunk) + ", 2^31)");
    g_msdMinKeys = n;
}
int sortDigitBits() { return g_digitBits; }
void setSortDigitBits(int bits)
{
    if (bits != 8 && bits != kWideBits) throw Error(1, "sort_digit_bits must be 8 or 11");
    g_digitBits = bits;
}
int sortDigitArray() { return g_digitArray; }
void setSortDigitArray(int on)
{
    if (on != 0 && on != 1) throw Error(1, "sort_digit_array must be 0 or 1");
    g_digitArray = on;
}
int sortTargetRanges() { return g_targetRanges; }
void setSortTargetRanges(int ranges)
{
    if (ranges < 64 || ranges > 65536) throw Error(1, "sort target ranges must be in [64, 65536]");
    g_targetRanges = ranges;
}

SortPlan makeSortPlan(std::size_t n)
{
    SortPlan p;
    p.n = n;
    p.items = g_items;
    p.threads = g_threads;
    if (p.items * p.threads > kMaxTileKeys)
        throw Error(1, "sort tile (sort_threads x sort_items) must be at most " + std::to_string(kMaxTileKeys) +
                           " keys; got " + std::to_string(p.threads) + " x " + std::to_string(p.items));
    p.tileKeys = std::size_t(p.items) * std::size_t(p.threads);
    const std::size_t tiles = (n + p.tileKeys - 1) / p.tileKeys;
    p.rangeTiles = (tiles + std::size_t(g_targetRanges) - 1) / std::size_t(g_targetRanges);
    if (p.rangeTiles < 1) p.rangeTiles = 1;
    p.rangeKeys = p.rangeTiles * p.tileKeys;
    p.ranges = (n + p.rangeKeys - 1) / p.rangeKeys;
    if (p.ranges < 1) p.ranges = 1;
    // enough histogram blocks to fill the GPU: parts per range, each a
    // multiple of 4 keys
    p.parts = 1;
    while (p.ranges * std::size_t(p.parts) < kHistogramBlocks && p.parts < 16 &&
           (p.rangeKeys / std::size_t(p.parts * 2)) % 4 == 0 && p.rangeKeys / std::size_t(p.parts * 2) >= 1024)
        p.parts *= 2;
    return p;
}

void SortWorkspace::allocate(std::size_t cap, bool withValues, hipStream_t s)
{
    if (cap <= capacity && (!withValues || values)) return;
    releaseOn(s);
    capacity = cap < 1 ? 1 : cap;
    // the smallest tile gives the largest range count for any tuning
    maxRanges = (capacity + 8 * kSortThreads - 1) / (8 * kSortThreads);
    if (maxRanges < 1) maxRanges = 1;
    const std::size_t histRows = maxRanges > kHistogramBlocks * 2 ? maxRanges : kHistogramBlocks * 2;
    values = withValues;
    keysAlt = static_cast<Key *>(poolAlloc(capacity * sizeof(Key), s));
    if (withValues) valsAlt = static_cast<std::uint32_t *>(poolAlloc(capacity * sizeof(std::uint32_t), s));
    digits = static_cast<std::uint8_t *>(poolAlloc((capacity + 16) * sizeof(std::uint8_t), s));
    rangeHist = static_cast<std::uint32_t *>(poolAlloc(histRows * kRadix * sizeof(std::uint32_t), s));
    rangeOff = static_cast<std::uint32_t *>(poolAlloc(maxRanges * kRadix * sizeof(std::uint32_t), s));
    digitTotal = static_cast<std::uint32_t *>(poolAlloc(kRadix * sizeof(std::uint32_t), s));
    if (capacity < kWideMaxKeys)
    {
        wideRanges = (capacity + kWideTile - 1) / kWideTile;
        wideHist = static_cast<std::uint32_t *>(poolAlloc(wideRanges * kWideRadix * sizeof(std::uint32_t), s));
        wideOff = static_cast<std::uint32_t *>(poolAlloc(wideRanges * kWideRadix * sizeof(std::uint32_t), s));
        wideTotal = static_cast<std::uint32_t *>(poolAlloc(kWideRadix * sizeof(std::uint32_t), s));
    }
}

void SortWorkspace::allocateMsd(std::size_t cap, std::size_t parents, hipStream_t s)
{
    if (cap <= msdCapacity && parents <= msdParents) return;
    for (void *p : {(void *)msdCnt, (void *)msdRoot, (void *)msdSegs[0], (void *)msdSegs[1], (void *)msdRanges[0],
                    (void *)msdRanges[1], (void *)msdFirst[0], (void *)msdFirst[1], (void *)msdHist, (void *)msdOff,
                    (void *)msdSubTotals, (void *)msdChunks})
        poolFree(p, s);
    msdCapacity = cap;
    msdParents = parents;
    // segments hold more than kMsdChunk keys; a level's ranges are its
    // segments' tiles; chunks: below 2 per kMsdChunk keys of each run of
    // small buckets, plus one run per parent and per segment over all levels
    // (level 1's parents: the whole array, or each segment of a segmented sort)
    msdMaxSegs = cap / kMsdChunk + 1;
    msdStride = ((cap / kMsdTile + msdMaxSegs + kUpRows) + 7) & ~std::size_t(7);
    msdMaxChunks = 16 * (cap / kMsdChunk) + 64 + parents;
    auto alloc = [&](std::size_t bytes) { return poolAlloc(bytes, s); };
    msdCnt = static_cast<MsdCounters *>(alloc(sizeof(MsdCounters)));
    if (!msdHostCount) msdHostCount = takePinnedWord();
    msdRoot = static_cast<MsdSeg *>(alloc(sizeof(MsdSeg)));
    for (int k = 0; k < 2; ++k)
    {
        msdSegs[k] = static_cast<MsdSeg *>(alloc(msdMaxSegs * sizeof(MsdSeg)));
        msdRanges[k] = static_cast<SegRange *>(alloc(msdStride * sizeof(SegRange)));
        msdFirst[k] = static_cast<std::uint8_t *>(alloc(msdStride));
    }
    msdHist = static_cast<std::uint32_t *>(alloc(std::size_t(kRadix) * msdStride * sizeof(std::uint32_t)));
    msdOff = static_cast<std::uint32_t *>(alloc(std::size_t(kRadix) * msdStride * sizeof(std::uint32_t)));
    msdSubTotals = static_cast<std::uint32_t *>(alloc(msdMaxSegs * kRadix * sizeof(std::uint32_t)));
    msdChunks = static_cast<MsdChunk *>(alloc(msdMaxChunks * sizeof(MsdChunk)));
}

// Frees every buffer through `give` (stream-ordered or after a
// synchronisation); the pinned count word is kept across regrowth (the host
// may poll it until the stream's work ends) and returned at destruction.
template <class Give>
void releaseWith(SortWorkspace &w, Give give, bool final)
{
    if (w.borrowed)
    {
        w.keysAlt = nullptr; w.valsAlt = nullptr; w.digits = nullptr; w.rangeHist = nullptr; w.rangeOff = nullptr;
        w.digitTotal = nullptr; w.wideHist = w.wideOff = w.wideTotal = nullptr; w.wideRanges = 0;
        w.capacity = 0; w.maxRanges = 0; w.values = false; w.borrowed = false;
        return;
    }
    for (void *p : {(void *)w.keysAlt, (void *)w.valsAlt, (void *)w.digits, (void *)w.rangeHist, (void *)w.rangeOff,
                    (void *)w.digitTotal, (void *)w.wideHist, (void *)w.wideOff, (void *)w.wideTotal})
        give(p);
    w.keysAlt = nullptr; w.valsAlt = nullptr; w.digits = nullptr; w.rangeHist = nullptr; w.rangeOff = nullptr;
    w.digitTotal = nullptr; w.wideHist = w.wideOff = w.wideTotal = nullptr; w.wideRanges = 0;
    w.capacity = 0; w.maxRanges = 0; w.values = false;
    for (void *p : {(void *)w.msdCnt, (void *)w.msdRoot, (void *)w.msdSegs[0], (void *)w.msdSegs[1],
                    (void *)w.msdRanges[0], (void *)w.msdRanges[1], (void *)w.msdFirst[0], (void *)w.msdFirst[1],
                    (void *)w.msdHist, (void *)w.msdOff, (void *)w.msdSubTotals, (void *)w.msdChunks})
        give(p);
    w.msdCnt = nullptr; w.msdRoot = nullptr; w.msdSegs[0] = w.msdSegs[1] = nullptr;
    w.msdRanges[0] = w.msdRanges[1] = nullptr; w.msdFirst[0] = w.msdFirst[1] = nullptr;
    w.msdHist = w.msdOff = w.msdSubTotals = nullptr; w.msdChunks = nullptr;
    w.msdCapacity = w.msdMaxSegs = w.msdStride = w.msdMaxChunks = w.msdParents = 0;
    if (final)
    {
        givePinnedWord(w.msdHostCount);
        w.msdHostCount = nullptr;
    }
}

void SortWorkspace::release()
{
    releaseWith(*this, [](void *p) { poolRelease(p); }, true);
}
void SortWorkspace::releaseOn(hipStream_t s)
{
    releaseWith(*this, [s](void *p) { poolFree(p, s); }, false);
}

template <class T>
void keygenAndSort(hipStream_t s, KeyKind kind, const KeygenParams<T> &p, std::size_t n, Key *keysOut,
                   std::uint32_t *valsOut, SortWorkspace &ws)
{
    if (n == 0) return;
    if (n >= (std::size_t(1) << 32)) throw Error(3, "radix sort supports fewer than 2^32 keys per tree");
    ws.allocate(n, true, s);
    const SortPlan pl = makeSortPlan(n);
    // MSD: keys into the alternate buffer, histogram of the top digit
    const bool msd = g_msd && n >= g_msdMinKeys;
    Key *keys = msd ? ws.keysAlt : keysOut;
    const int histShift = msd ? 64 - kRadixBits : 0;
    {
        prof::Scope scope("keygen", s);
        const dim3 grid(unsigned(pl.ranges * pl.parts)), block(kSortThreads);
        switch (kind)
        {
        case KeyKind::Morton:
            hipLaunchKernelGGL((keygenKernel<T, KeyKind::Morton>), grid, block, 0, s, p, n, pl.rangeKeys, pl.parts,
                               keys, ws.rangeHist, histShift);
            break;
        case KeyKind::Hilbert:
            hipLaunchKernelGGL((keygenKernel<T, KeyKind::Hilbert>), grid, block, 0, s, p, n, pl.rangeKeys, pl.parts,
                               keys, ws.rangeHist, histShift);
            break;
        case KeyKind::Binary:
            hipLaunchKernelGGL((keygenKernel<T, KeyKind::Binary>), grid, block, 0, s, p, n, pl.rangeKeys, pl.parts,
                               keys, ws.rangeHist, histShift);
            break;
        }
    }
    if (msd) msdSort(s, pl, ws, ws.keysAlt, keysOut, ws.valsAlt, valsOut, true, false, true);
    else if (widePasses(ws, n)) runWidePasses(s, n, ws, keysOut, ws.keysAlt, valsOut, ws.valsAlt, true, true);
    else runPasses(s, pl, ws, keysOut, ws.keysAlt, valsOut, ws.valsAlt, true, true, false);
}

void sortPairs(hipStream_t s, Key *keys, std::uint32_t *vals, std::size_t n, SortWorkspace &ws, bool iota,
               bool msdAllowed)
{
    if (n == 0) return;
    if (n >= (std::size_t(1) << 32)) throw Error(3, "radix sort supports fewer than 2^32 keys");
    ws.allocate(n, true, s);
    const SortPlan pl = makeSortPlan(n);
    const bool msd = msdAllowed && g_msd && n >= g_msdMinKeys;
    if (!msd && widePasses(ws, n))
    {
        runWidePasses(s, n, ws, keys, ws.keysAlt, vals, ws.valsAlt, true, iota);
        return;
    }
    hipLaunchKernelGGL(upsweepKeysKernel, dim3(unsigned(pl.ranges)), dim3(kSortThreads), 0, s, keys, n, pl.rangeKeys,
                       msd ? 64 - kRadixBits : 0, ws.rangeHist);
    if (msd) msdSort(s, pl, ws, keys, ws.keysAlt, vals, ws.valsAlt, iota, true, false);
    else runPasses(s, pl, ws, keys, ws.keysAlt, vals, ws.valsAlt, true, iota, true);
}

void sortKeys(hipStream_t s, Key *keys, std::size_t n, SortWorkspace &ws)
{
    if (n == 0) return;
    if (n >= (std::size_t(1) << 32)) throw Error(3, "radix sort supports fewer than 2^32 keys");
    ws.allocate(n, false, s);
    if (widePasses(ws, n))
    {
        runWidePasses(s, n, ws, keys, ws.keysAlt, nullptr, nullptr, false, false);
        return;
    }
    const SortPlan pl = makeSortPlan(n);
    hipLaunchKernelGGL(upsweepKeysKernel, dim3(unsigned(pl.ranges)), dim3(kSortThreads), 0, s, keys, n, pl.rangeKeys,
                       0, ws.rangeHist);
    runPasses(s, pl, ws, keys, ws.keysAlt, nullptr, nullptr, false, false, true);
}

// ---- sort without allocation ------------------------------------------------------
// The LSD path of sortPairs over a workspace carved from `scratch`. Its
// largest plan: kNoAllocMaxKeys keys in tiles of the smallest shape (256 x 8
// keys) gives the most ranges, each with one histogram row of 256 digits.
namespace
{
constexpr std::size_t naAlign(std::size_t b) { return (b + 255) & ~std::size_t(255); }
constexpr std::size_t kNaRanges = (kNoAllocMaxKeys + 8 * kSortThreads - 1) / (8 * kSortThreads);
constexpr std::size_t kNaKeys = 0;
constexpr std::size_t kNaVals = kNaKeys + naAlign(kNoAllocMaxKeys * sizeof(Key));
constexpr std::size_t kNaDigits = kNaVals + naAlign(kNoAllocMaxKeys * sizeof(std::uint32_t));
constexpr std::size_t kNaHist = kNaDigits + naAlign(kNoAllocMaxKeys + 16);
constexpr std::size_t kNaOff = kNaHist + naAlign(kNaRanges * kRadix * sizeof(std::uint32_t));
constexpr std::size_t kNaTotal = kNaOff + naAlign(kNaRanges * kRadix * sizeof(std::uint32_t));
constexpr std::size_t kNaBytes = kNaTotal + naAlign(kRadix * sizeof(std::uint32_t));
} // namespace

// 14.1 MiB per GPU. Host code never uses this name as an address: the
// device address of each GPU's copy comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace
{
constexpr int kNaMaxDevices = 64;
std::mutex g_naMu;
unsigned char *g_naBase[kNaMaxDevices] = {};

// Device address of the current GPU's scratch array, looked up once per
// device with that device current.
unsigned char *noAllocBase()
{
    int dev = 0;
    AT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaMaxDevices) throw Error(1, "device index out of range for the no-alloc sort");
    std::lock_guard<std::mutex> lk(g_naMu);
    if (!g_naBase[dev])
    {
        void *ptr = nullptr;
        AT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw Error(4, "hipGetSymbolAddress returned no address for the sort scratch array");
        g_naBase[dev] = static_cast<unsigned char *>(ptr);
    }
    return g_naBase[dev];
}
} // namespace

std::size_t noAllocScratchBytes() { return kNaBytes; }

void sortPairsNoAlloc(hipStream_t s, Key *keys, std::uint32_t *vals, std::size_t n, bool iota)
{
    if (n > kNoAllocMaxKeys)
        throw Error(3, "sort_pairs_noalloc sorts at most " + std::to_string(kNoAllocMaxKeys) + " pairs; got " +
                           std::to_string(n));
    if (n == 0) return;
    unsigned char *base = noAllocBase();
    const SortPlan pl = makeSortPlan(n);
    if (pl.ranges > kNaRanges) throw Error(1, "sort plan exceeds the no-alloc workspace");
    SortWorkspace ws;
    ws.borrowed = true;
    ws.keysAlt = reinterpret_cast<Key *>(base + kNaKeys);
    ws.valsAlt = reinterpret_cast<std::uint32_t *>(base + kNaVals);
    ws.digits = base + kNaDigits;
    ws.rangeHist = reinterpret_cast<std::uint32_t *>(base + kNaHist);
    ws.rangeOff = reinterpret_cast<std::uint32_t *>(base + kNaOff);
    ws.digitTotal = reinterpret_cast<std::uint32_t *>(base + kNaTotal);
    ws.capacity = kNoAllocMaxKeys;
    ws.maxRanges = kNaRanges;
    ws.values = true;
    // every histogram row and offset the passes read is written first in the
    // same call, so nothing needs clearing
    hipLaunchKernelGGL(upsweepKeysKernel, dim3(unsigned(pl.ranges)), dim3(kSortThreads), 0, s, keys, n, pl.rangeKeys,
                       0, ws.rangeHist);
    runPasses(s, pl, ws, keys, ws.keysAlt, vals, ws.valsAlt, true, iota, true, false);
}

// ---- autotuner ------------------------------------------------------------------
namespace
{
// Keys with full 64-bit entropy from a counter hash (splitmix64 finaliser).
__global__ void fillTuneKeysKernel(Key *keys, std::size_t n, std::uint64_t seed)
{
    for (std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += std::size_t(gridDim.x) * blockDim.x)
    {
        std::uint64_t z = seed + (i + 1) * 0x9e3779b97f4a7c15ull;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        keys[i] = (z ^ (z >> 31)) >> 1; // 63-bit keys, like every SFC key kind
    }
}
} // namespace

std::string autotuneSort(std::size_t n, hipStream_t s)
{
    if (n < 1024) n = 1024;
    if (n >= (std::size_t(1) << 32)) throw Error(1, "autotune size must be below 2^32 keys");
    DevBuf<Key> src, keys;
    DevBuf<std::uint32_t> vals;
    src.ensure(n, 0, s);
    keys.ensure(n, 0, s);
    vals.ensure(n, 0, s);
    SortWorkspace ws;
    ws.allocate(n, true, s);
    hipLaunchKernelGGL(fillTuneKeysKernel, dim3(2048), dim3(256), 0, s, src.p, n, 0x5eedull);
    hipEvent_t e0, e1;
    AT_HIP(hipEventCreate(&e0));
    AT_HIP(hipEventCreate(&e1));
    auto timeOne = [&]() {
        float best = 1e30f;
        for (int rep = 0; rep < 4; ++rep)
        {
            AT_HIP(hipMemcpyAsync(keys.p, src.p, n * sizeof(Key), hipMemcpyDeviceToDevice, s));
            AT_HIP(hipEventRecord(e0, s));
            sortPairs(s, keys.p, vals.p, n, ws, true);
            AT_HIP(hipEventRecord(e1, s));
            AT_HIP(hipEventSynchronize(e1));
            float ms = 0;
            AT_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0 && ms < best) best = ms; // the first run warms up
        }
        return best;
    };
    // coordinate descent: tile shape (with / without next-tile prefetch),
    // then ranges, then the histogram source
    const int shapes[][3] = {{256, 8, 0},  {256, 12, 0}, {256, 16, 0}, {512, 8, 0},  {512, 12, 0}, {512, 16, 0},
                             {1024, 8, 0}, {256, 8, 1}, {256, 12, 1}, {512, 8, 1}, {512, 12, 1}};
    float best = 1e30f;
    int bt = g_threads, bi = g_items, bp = g_prefetch;
    for (const auto &sh : shapes)
    {
        g_threads = sh[0];
        g_items = sh[1];
        g_prefetch = sh[2];
        const float t = timeOne();
        if (t < best) best = t, bt = sh[0], bi = sh[1], bp = sh[2];
    }
    g_threads = bt;
    g_items = bi;
    g_prefetch = bp;
    int br = g_targetRanges;
    for (int r : {2048, 4096, 8192, 16384, 32768})
    {
        g_targetRanges = r;
        const float t = timeOne();
        if (t < best) best = t, br = r;
    }
    g_targetRanges = br;
    int bd = g_digitArray;
    for (int d : {0, 1})
    {
        g_digitArray = d;
        const float t = timeOne();
        if (t < best) best = t, bd = d;
    }
    g_digitArray = bd;
    int bm = g_msd;
    if (n >= g_msdMinKeys)
        for (int m : {0, 1})
        {
            g_msd = m;
            const float t = timeOne();
            if (t < best) best = t, bm = m;
        }
    g_msd = bm;
    // the LSD digit width, on a sort below the MSD threshold (its own
    // workspace: the 11-bit histograms exist only for such sizes)
    const std::size_t m = n < kWideMaxKeys ? n : kWideMaxKeys - 1;
    SortWorkspace small;
    small.allocate(m, true, s);
    auto timeLsd = [&]() {
        float b = 1e30f;
        for (int rep = 0; rep < 6; ++rep)
        {
            AT_HIP(hipMemcpyAsync(keys.p, src.p, m * sizeof(Key), hipMemcpyDeviceToDevice, s));
            AT_HIP(hipEventRecord(e0, s));
            sortPairs(s, keys.p, vals.p, m, small, true, false);
            AT_HIP(hipEventRecord(e1, s));
            AT_HIP(hipEventSynchronize(e1));
            float ms = 0;
            AT_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0 && ms < b) b = ms;
        }
        return b;
    };
    float lsd[2] = {0.f, 0.f};
    int bb = 8;
    float bestLsd = 1e30f;
    for (int i = 0; i < 2; ++i)
    {
        g_digitBits = i ? kWideBits : 8;
        lsd[i] = timeLsd();
        if (lsd[i] < bestLsd) bestLsd = lsd[i], bb = g_digitBits;
    }
    g_digitBits = bb;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    syncStream(s);
    return "{\"sort_threads\": " + std::to_string(bt) + ", \"sort_items\": " + std::to_string(bi) +
           ", \"sort_prefetch\": " + std::to_string(bp) +
           ", \"sort_ranges\": " + std::to_string(br) + ", \"sort_digit_array\": " + std::to_string(bd) +
           ", \"sort_msd\": " + std::to_string(bm) + ", \"sort_msd_min_keys\": " + std::to_string(g_msdMinKeys) +
           ", \"keys\": " + std::to_string(n) + ", \"sort_ms\": " + std::to_string(best) +
           ", \"sort_digit_bits\": " + std::to_string(bb) + ", \"digit_bits_keys\": " + std::to_string(m) +
           ", \"lsd_ms_8_bits\": " + std::to_string(lsd[0]) + ", \"lsd_ms_11_bits\": " + std::to_string(lsd[1]) +
           "}";
}

// ---- SegmentedSort -------------------------------------------------------------
void SegmentedSort::plan(const std::vector<std::size_t> &counts, hipStream_t s)
{
    segments_ = counts.size();
    segBegin_.assign(segments_ + 1, 0);
    for (std::size_t i = 0; i < segments_; ++i) segBegin_[i + 1] = segBegin_[i] + counts[i];
    total_ = segBegin_[segments_];
    if (total_ >= (std::size_t(1) << 32)) throw Error(3, "segmented sort supports fewer than 2^32 keys in total");
    constexpr std::size_t tile = std::size_t(kSegThreads) * kSegItems;
    rangeKeys_ = tile; // one tile per range (see the block-to-range mapping of scatterKernel)
    std::vector<SegRange> rs;
    std::vector<std::uint32_t> first(segments_ + 1, 0);
    for (std::size_t g = 0; g < segments_; ++g)
    {
        first[g] = std::uint32_t(rs.size());
        for (std::size_t b = 0; b < counts[g]; b += rangeKeys_)
        {
            const std::size_t len = counts[g] - b < rangeKeys_ ? counts[g] - b : rangeKeys_;
            rs.push_back(SegRange{std::uint64_t(segBegin_[g] + b), std::uint32_t(len), std::uint32_t(b),
                                  std::uint32_t(g), 0u});
        }
    }
    first[segments_] = std::uint32_t(rs.size());
    ranges_ = rs.size();
    std::vector<std::uint64_t> begins(segBegin_.begin(), segBegin_.end());
    range_.ensure(ranges_ ? ranges_ : 1, 0, s);
    segRange_.ensure(segments_ + 1, 0, s);
    segBeginD_.ensure(segments_ + 1, 0, s);
    hist_.ensure((ranges_ ? ranges_ : 1) * kRadix, 0, s);
    off_.ensure((ranges_ ? ranges_ : 1) * kRadix, 0, s);
    keysAlt_.ensure(total_ ? total_ : 1, 0, s);
    valsAlt_.ensure(total_ ? total_ : 1, 0, s);
    digits_.ensure(total_ ? total_ : 1, 0, s);
    msdReady_ = total_ >= g_msdMinKeys;
    std::vector<MsdSeg> parents;
    if (msdReady_)
    {
        msd_.allocateMsd(total_, segments_, s);
        parents_.ensure(segments_, 0, s);
        parentCount_.ensure(1, 0, s);
        segTotals_.ensure(segments_ * kRadix, 0, s);
        for (std::size_t g = 0; g < segments_; ++g)
            parents.push_back(MsdSeg{std::uint64_t(segBegin_[g]), std::uint32_t(counts[g]), 0, 0, 0});
    }
    syncStream(s);
    if (ranges_) AT_HIP(hipMemcpy(range_.p, rs.data(), ranges_ * sizeof(SegRange), hipMemcpyHostToDevice));
    AT_HIP(hipMemcpy(segRange_.p, first.data(), first.size() * sizeof(std::uint32_t), hipMemcpyHostToDevice));
    AT_HIP(hipMemcpy(segBeginD_.p, begins.data(), begins.size() * sizeof(std::uint64_t), hipMemcpyHostToDevice));
    if (msdReady_)
    {
        const std::uint32_t nseg = std::uint32_t(segments_);
        AT_HIP(hipMemcpy(parents_.p, parents.data(), segments_ * sizeof(MsdSeg), hipMemcpyHostToDevice));
        AT_HIP(hipMemcpy(parentCount_.p, &nseg, sizeof(nseg), hipMemcpyHostToDevice));
    }
}

template <class T>
void SegmentedSort::keygenAndSort(hipStream_t s, KeyKind kind, const KeygenParams<T> &shared,
                                  const SegCoords<T> *coords, Key *keys, std::uint32_t *vals, bool hostPoll)
{
    if (!ranges_) return;
    const dim3 rgrid{unsigned(ranges_)}, block{unsigned(kSortThreads)};
    // MSD: keys into the alternate buffer, histograms of the top digit
    const bool msd = g_msd && msdReady_;
    Key *kgen = msd ? keysAlt_.p : keys;
    const int histShift = msd ? 64 - kRadixBits : 0;
    {
        prof::Scope scope("keygen", s);
        switch (kind)
        {
        case KeyKind::Morton:
            hipLaunchKernelGGL((segKeygenKernel<T, KeyKind::Morton>), rgrid, block, 0, s, shared, coords, range_.p,
                               kgen, hist_.p, histShift);
            break;
        case KeyKind::Hilbert:
            hipLaunchKernelGGL((segKeygenKernel<T, KeyKind::Hilbert>), rgrid, block, 0, s, shared, coords, range_.p,
                               kgen, hist_.p, histShift);
            break;
        case KeyKind::Binary:
            hipLaunchKernelGGL((segKeygenKernel<T, KeyKind::Binary>), rgrid, block, 0, s, shared, coords, range_.p,
                               kgen, hist_.p, histShift);
            break;
        }
    }
    if (msd)
    {
        // level 1 per segment (slot values), then the segments are the parents
        // of the MSD levels; the result ends in keys / vals
        hipLaunchKernelGGL(msdResetKernel, dim3(1), dim3(1), 0, s, msd_.msdCnt, msd_.msdRoot, total_);
        hipLaunchKernelGGL(segScanKernel, dim3(unsigned(segments_)), block, 0, s, hist_.p, segRange_.p,
                           segBeginD_.p, off_.p, segTotals_.p);
        {
            prof::Scope scope("sort_pass", s);
            hipLaunchKernelGGL((scatterKernel<kSegThreads, kSegItems, true, true, true, true>), dim3(unsigned(ranges_)),
                               dim3(unsigned(kSegThreads)), 0, s, keysAlt_.p, keys, valsAlt_.p, vals, digits_.p, total_,
                               rangeKeys_, off_.p, nullptr, ranges_, 64 - kRadixBits, 64 - 2 * kRadixBits, range_.p);
        }
        hipLaunchKernelGGL(msdPlanKernel, dim3(unsigned(segments_ < 1024 ? segments_ : 1024)), dim3(kSortThreads), 0,
                           s, 1, parents_.p, parentCount_.p, segTotals_.p, msd_.msdCnt, 0, msd_.msdSegs[0],
                           msd_.msdRanges[0], msd_.msdFirst[0], std::uint32_t(msd_.msdMaxSegs),
                           std::uint32_t(msd_.msdStride), msd_.msdChunks, std::uint32_t(msd_.msdMaxChunks),
                           std::uint32_t(segments_));
        msdLevels(s, msd_, total_, digits_.p, keysAlt_.p, keys, valsAlt_.p, vals, true, hostPoll);
        return;
    }
    Key *kA = keys, *kB = keysAlt_.p;
    std::uint32_t *vA = vals, *vB = valsAlt_.p;
    for (int pass = 0; pass < kSortPasses; ++pass)
    {
        const Key *src = (pass & 1) ? kB : kA;
        Key *dst = (pass & 1) ? kA : kB;
        const std::uint32_t *vsrc = (pass & 1) ? vB : vA;
        std::uint32_t *vdst = (pass & 1) ? vA : vB;
        const int shift = pass * kRadixBits, next = shift + kRadixBits;
        if (pass > 0)
        {
            prof::Scope scope("sort_upsweep", s);
            hipLaunchKernelGGL(segUpsweepDigitsKernel, rgrid, block, 0, s, digits_.p, range_.p, hist_.p);
        }
        hipLaunchKernelGGL(segScanKernel, dim3(unsigned(segments_)), block, 0, s, hist_.p, segRange_.p,
                           segBeginD_.p, off_.p, nullptr);
        prof::Scope scope("sort_pass", s);
        const dim3 sgrid{unsigned(ranges_)}, sblock{unsigned(kSegThreads)};
#define AT_SEG_SCATTER(I, D)                                                                                           \
    hipLaunchKernelGGL((scatterKernel<kSegThreads, kSegItems, true, I, D, true>), sgrid, sblock, 0, s, src, dst, vsrc, \
                       vdst, digits_.p, total_, rangeKeys_, off_.p, nullptr, ranges_, shift, next, range_.p)
        if (pass == 0) AT_SEG_SCATTER(true, true);
        else if (pass + 1 < kSortPasses) AT_SEG_SCATTER(false, true);
        else AT_SEG_SCATTER(false, false);
#undef AT_SEG_SCATTER
    }
    AT_HIP(hipGetLastError());
}

template void SegmentedSort::keygenAndSort<float>(hipStream_t, KeyKind, const KeygenParams<float> &,
                                                  const SegCoords<float> *, Key *, std::uint32_t *, bool);
template void SegmentedSort::keygenAndSort<double>(hipStream_t, KeyKind, const KeygenParams<double> &,
                                                   const SegCoords<double> *, Key *, std::uint32_t *, bool);

template void keygenAndSort<float>(hipStream_t, KeyKind, const KeygenParams<float> &, std::size_t, Key *,
                                   std::uint32_t *, SortWorkspace &);
template void keygenAndSort<double>(hipStream_t, KeyKind, const KeygenParams<double> &, std::size_t, Key *,
                                    std::uint32_t *, SortWorkspace &);
} // namespace at
