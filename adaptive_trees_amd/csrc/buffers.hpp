// buffers.hpp -- device allocations and thrust-style capacity bookkeeping.
#pragma once

#include <utility>

#include "common.hpp"

namespace at
{
// Tracks the logical size and allocation capacity of one of the reference's
// thrust::device_vector buffers with thrust's growth rules (reserve(n)
// allocates exactly n; a growing resize allocates max(2*size, n,
// 2*capacity)). The octree's leafCapacity is derived from these capacities
// after build (octree.cu:453-457), so they are mirrored exactly; the device
// memory itself is managed separately.
struct VecShape
{
    std::size_t size = 0, cap = 0;
    void reserve(std::size_t n)
    {
        if (n > cap) cap = n;
    }
    void resize(std::size_t n)
    {
        if (n > cap)
        {
            std::size_t c = 2 * size > n ? 2 * size : n;
            cap = c > 2 * cap ? c : 2 * cap;
        }
        size = n;
    }
    void resizeHeadroom(std::size_t req) // resizeWithHeadroom (helpers.cuh:122-141)
    {
        if (req > cap) reserve(req + req / 4);
        resize(req);
    }
};

// Device memory of the engine comes from the device's default stream-ordered
// pool (its release threshold raised once per process), in size classes.
// Allocating and freeing never synchronises the device (hipFree does),
// which matters when trees are created and released while others compute.
//   poolFree(p, s)  p may still be used by work queued on s: it is kept
//                   pending and becomes reusable when the library next
//                   synchronises s (syncStream);
//   poolRelease(p)  the owner has synchronised every stream that used p:
//                   reusable at once, by the next allocation of its size
//                   class, without a runtime call.
void *poolAlloc(std::size_t bytes, hipStream_t s);
void poolFree(void *p, hipStream_t s);
void poolRelease(void *p);
// hipStreamSynchronize, after which the blocks freed on s before the call
// are reusable. syncStream throws on a HIP error; syncStreamStatus returns it.
void syncStream(hipStream_t s);
hipError_t syncStreamStatus(hipStream_t s);
// A word of pinned, coherent host memory (device stores are visible to a
// polling host), from slabs kept until the process ends.
std::uint32_t *takePinnedWord();
void givePinnedWord(std::uint32_t *w);

// A device allocation that grows on request.
template <class T>
struct DevBuf
{
    T *p = nullptr;
    std::size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    // the owner has synchronised its stream (destruction)
    ~DevBuf() { release(); }
    void release()
    {
        poolRelease(p);
        p = nullptr;
        cap = 0;
    }
    // Grow to at least n elements; the first `keep` elements survive.
    void ensure(std::size_t n, std::size_t keep, hipStream_t s)
    {
        if (n <= cap) return;
        T *q = static_cast<T *>(poolAlloc((n ? n : 1) * sizeof(T), s));
        if (p && keep)
            AT_HIP(hipMemcpyAsync(q, p, (keep < cap ? keep : cap) * sizeof(T), hipMemcpyDeviceToDevice, s));
        poolFree(p, s); // reusable once s has run past the copy and every earlier use
        p = q;
        cap = n;
    }
    void swap(DevBuf &o)
    {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
    }
};

} // namespace at
