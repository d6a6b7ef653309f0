// engine.hip -- tree construction entry, shared state helpers and the C ABI.
#include "engine.hpp"
#include "tree_impl.hpp"

#include <map>
#include <mutex>
#include <vector>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>

namespace at
{
void hipCheck(hipError_t e, const char *what, const char *file, int line)
{
    if (e != hipSuccess)
        throw Error(AT_EHIP, std::string(what) + " failed at " + file + ":" + std::to_string(line) + ": " +
                                 hipGetErrorString(e));
}

namespace
{
std::once_flag g_poolOnce[64];

// Idle device blocks kept for reuse. Blocks come from the device's
// stream-ordered pool in size classes (8 per power of two, at most 12.5%
// over the request) and stay owned by this cache:
//   idle      free for the next request of their class on their device,
//             taken without a runtime call (up to kIdleCap bytes per device);
//   pending   freed in the order of a stream whose work may still use them:
//             they become idle when the library next synchronises that
//             stream (syncStream), having taken the list before the wait.
// Trees are created and destroyed by the thousand in a serving loop and
// grow buffers pass by pass in a build: hipMallocAsync costs a few
// microseconds and hipFreeAsync tens (it may trim the pool), per buffer.
class BlockCache
{
public:
    static constexpr std::size_t kIdleCap = std::size_t(16) << 30;
    static constexpr std::size_t kPendingCap = std::size_t(4) << 30; // per stream; beyond: hipFreeAsync
    static std::size_t classOf(std::size_t bytes)
    {
        if (bytes <= 256) return 256;
        std::size_t p = std::size_t(1) << (63 - __builtin_clzll(bytes)); // highest power of two <= bytes
        const std::size_t step = p / 8;
        return (bytes + step - 1) / step * step;
    }
    void *take(int dev, std::size_t cls)
    {
        std::lock_guard<std::mutex> g(m_);
        auto it = idle_.find({dev, cls});
        if (it == idle_.end() || it->second.empty()) return nullptr;
        void *p = it->second.back();
        it->second.pop_back();
        idleBytes_[dev & 63] -= cls;
        return p;
    }
    void own(void *p, int dev, std::size_t cls)
    {
        std::lock_guard<std::mutex> g(m_);
        owned_[p] = {dev, cls};
    }
    // The owner has synchronised: true = kept idle; false = the caller frees
    // it (forgotten here).
    bool release(void *p)
    {
        std::lock_guard<std::mutex> g(m_);
        return idleLocked(p);
    }
    // Freed in stream s's order: true = pending until s is synchronised;
    // false = the caller frees it on s (forgotten here).
    bool defer(void *p, hipStream_t s)
    {
        std::lock_guard<std::mutex> g(m_);
        auto it = owned_.find(p);
        if (it == owned_.end()) return false;
        Pending &q = pending_[s];
        if (q.bytes + it->second.second > kPendingCap)
        {
            owned_.erase(it);
            return false;
        }
        q.blocks.push_back(p);
        q.bytes += it->second.second;
        return true;
    }
    std::vector<void *> takePending(hipStream_t s)
    {
        std::lock_guard<std::mutex> g(m_);
        auto it = pending_.find(s);
        if (it == pending_.end()) return {};
        std::vector<void *> out;
        out.swap(it->second.blocks);
        pending_.erase(it);
        return out;
    }
    // Blocks whose stream has been synchronised: idle; those beyond the idle
    // cap are returned for the caller to free.
    std::vector<void *> settle(const std::vector<void *> &blocks)
    {
        std::lock_guard<std::mutex> g(m_);
        std::vector<void *> drop;
        for (void *p : blocks)
            if (!idleLocked(p)) drop.push_back(p);
        return drop;
    }

private:
    bool idleLocked(void *p)
    {
        auto it = owned_.find(p);
        if (it == owned_.end()) return false;
        const auto [dev, cls] = it->second;
        if (idleBytes_[dev & 63] + cls <= kIdleCap)
        {
            idle_[{dev, cls}].push_back(p);
            idleBytes_[dev & 63] += cls;
            return true;
        }
        owned_.erase(it);
        return false;
    }
    struct Pending
    {
        std::vector<void *> blocks;
        std::size_t bytes = 0;
    };
    std::mutex m_;
    std::map<void *, std::pair<int, std::size_t>> owned_;
    std::map<std::pair<int, std::size_t>, std::vector<void *>> idle_;
    std::map<hipStream_t, Pending> pending_;
    std::size_t idleBytes_[64] = {};
};
BlockCache &blockCache()
{
    static BlockCache *c = new BlockCache; // never destroyed: blocks stay valid at exit
    return *c;
}

// Idle blocks beyond the cache's cap are freed on a non-blocking stream of
// their device (never the legacy default stream: while any stream of the
// process is being captured, HIP fails legacy-stream work in every thread).
hipStream_t g_reclaim[64] = {};

void setupDevice(int dev)
{
    std::call_once(g_poolOnce[dev & 63], [dev] {
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool)
        {
            std::uint64_t threshold = ~std::uint64_t(0);
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold);
        }
        (void)hipStreamCreateWithFlags(&g_reclaim[dev & 63], hipStreamNonBlocking);
    });
}
hipStream_t reclaimStream(void *p)
{
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) return nullptr;
    return g_reclaim[a.device & 63];
}
} // namespace

void *poolAlloc(std::size_t bytes, hipStream_t s)
{
    int dev = 0;
    AT_HIP(hipGetDevice(&dev));
    setupDevice(dev);
    const std::size_t cls = BlockCache::classOf(bytes ? bytes : 1);
    if (void *p = blockCache().take(dev, cls)) return p;
    void *p = nullptr;
    AT_HIP(hipMallocAsync(&p, cls, s));
    blockCache().own(p, dev, cls);
    return p;
}

void poolFree(void *p, hipStream_t s)
{
    if (!p) return;
    if (blockCache().defer(p, s)) return;
    (void)hipFreeAsync(p, s);
}

void poolRelease(void *p)
{
    if (!p) return;
    if (blockCache().release(p)) return;
    (void)hipFreeAsync(p, reclaimStream(p)); // idle: no stream uses it
}

namespace
{
hipError_t syncAndSettle(hipStream_t s)
{
    std::vector<void *> blocks = blockCache().takePending(s);
    const hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess)
    {
        // the stream's work may not have finished: its blocks are freed in its order
        for (void *p : blocks) (void)hipFreeAsync(p, s);
        return e;
    }
    for (void *p : blockCache().settle(blocks)) (void)hipFreeAsync(p, reclaimStream(p)); // idle now
    return hipSuccess;
}
} // namespace

void syncStream(hipStream_t s) { AT_HIP(syncAndSettle(s)); }
hipError_t syncStreamStatus(hipStream_t s) { return syncAndSettle(s); }

namespace
{
// Pinned host slots for the per-tree state read-backs, carved from slabs that
// live until the process ends: trees are created and destroyed without
// hipHostMalloc / hipHostFree (which synchronise the device) and read back
// through pinned memory.
class PinnedStates
{
public:
    DevState *take()
    {
        std::lock_guard<std::mutex> g(m_);
        if (free_.empty())
        {
            constexpr std::size_t kSlab = 1024;
            DevState *slab = nullptr;
            AT_HIP(hipHostMalloc(&slab, kSlab * sizeof(DevState)));
            for (std::size_t i = 0; i < kSlab; ++i) free_.push_back(slab + i);
        }
        DevState *s = free_.back();
        free_.pop_back();
        *s = DevState{};
        return s;
    }
    void give(DevState *s)
    {
        std::lock_guard<std::mutex> g(m_);
        free_.push_back(s);
    }

private:
    std::mutex m_;
    std::vector<DevState *> free_;
};
PinnedStates &pinnedStates()
{
    static PinnedStates *p = new PinnedStates; // never destroyed: slabs stay valid at exit
    return *p;
}

class PinnedWords
{
public:
    std::uint32_t *take()
    {
        std::lock_guard<std::mutex> g(m_);
        if (free_.empty())
        {
            constexpr std::size_t kSlab = 4096;
            std::uint32_t *slab = nullptr;
            AT_HIP(hipHostMalloc(reinterpret_cast<void **>(&slab), kSlab * sizeof(std::uint32_t),
                                 hipHostMallocCoherent));
            for (std::size_t i = 0; i < kSlab; ++i) free_.push_back(slab + i);
        }
        std::uint32_t *w = free_.back();
        free_.pop_back();
        return w;
    }
    void give(std::uint32_t *w)
    {
        std::lock_guard<std::mutex> g(m_);
        free_.push_back(w);
    }

private:
    std::mutex m_;
    std::vector<std::uint32_t *> free_;
};
PinnedWords &pinnedWords()
{
    static PinnedWords *p = new PinnedWords;
    return *p;
}
} // namespace

std::uint32_t *takePinnedWord() { return pinnedWords().take(); }
void givePinnedWord(std::uint32_t *w)
{
    if (w) pinnedWords().give(w);
}

TreeBase::TreeBase(hipStream_t s, DevState *st, DevState *sth) : stream_(s)
{
    if (st && sth)
    {
        st_ = st;
        sth_ = sth;
        ownsState_ = false;
        return;
    }
    st_ = static_cast<DevState *>(poolAlloc(sizeof(DevState), stream_));
    sth_ = pinnedStates().take();
    AT_HIP(hipMemsetAsync(st_, 0, sizeof(DevState), stream_));
}

TreeBase::~TreeBase()
{
    (void)syncStreamStatus(stream_);
    if (!ownsState_) return;
    poolRelease(st_);
    pinnedStates().give(sth_);
}

void TreeBase::readState()
{
    AT_HIP(hipMemcpyAsync(sth_, st_, sizeof(DevState), hipMemcpyDeviceToHost, stream_));
    syncStream(stream_);
}

using impl::makeTree;

TreeBase *createTree(int realBytes, int type, const void *x, const void *y, const void *z, std::size_t n,
                     const at_config &c)
{
    if (realBytes == 4)
        return makeTree<float>(type, static_cast<const float *>(x), static_cast<const float *>(y),
                               static_cast<const float *>(z), n, c);
    if (realBytes == 8)
        return makeTree<double>(type, static_cast<const double *>(x), static_cast<const double *>(y),
                                static_cast<const double *>(z), n, c);
    throw Error(AT_EINVAL, "real_bytes must be 4 or 8");
}
} // namespace at

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace
{
thread_local std::string g_error;
}
std::string &at::lastError() { return g_error; }
namespace
{

template <class F>
int guarded(F &&f)
{
    try
    {
        f();
        return AT_OK;
    }
    catch (const at::Error &e)
    {
        g_error = e.what();
        return e.code;
    }
    catch (const std::bad_alloc &e)
    {
        g_error = e.what();
        return AT_ENOMEM;
    }
    catch (const std::exception &e)
    {
        g_error = e.what();
        return AT_EINVAL;
    }
}

int hipStatus(hipError_t e)
{
    if (e == hipSuccess) return AT_OK;
    g_error = hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? AT_ENOMEM : AT_EHIP;
}
} // namespace

struct at_tree
{
    std::unique_ptr<at::TreeBase> impl;
};

namespace
{
at::TreeBase &treeOf(at_tree *t)
{
    if (!t || !t->impl) throw at::Error(AT_EINVAL, "null tree");
    return *t->impl;
}
template <class... P>
void needOutputs(P *...p)
{
    if (((p == nullptr) || ...)) throw at::Error(AT_EINVAL, "null output pointer");
}
} // namespace

extern "C" {
int at_tree_create(int realBytes, int type, const void *x, const void *y, const void *z, size_t count,
                   const at_config *config, at_tree **out)
{
    return guarded([&] {
        if (!config || !out) throw at::Error(AT_EINVAL, "null argument");
        auto t = std::make_unique<at_tree>();
        t->impl.reset(at::createTree(realBytes, type, x, y, z, count, *config));
        *out = t.release();
    });
}
int at_tree_destroy(at_tree *t)
{
    return guarded([&] { delete t; });
}
int at_tree_build(at_tree *t)
{
    return guarded([&] { treeOf(t).build(); });
}
int at_tree_update(at_tree *t)
{
    return guarded([&] { treeOf(t).update(); });
}
int at_tree_compute_views(at_tree *t, const uint8_t *v)
{
    return guarded([&] {
        if (v)
        {
            at::ViewFlags f{v[0] != 0, v[1] != 0, v[2] != 0};
            treeOf(t).computeViews(&f);
        }
        else
            treeOf(t).computeViews(nullptr);
    });
}
int at_tree_needs_resize(at_tree *t, int *o)
{
    return guarded([&] {
        needOutputs(o);
        *o = treeOf(t).needsResize() ? 1 : 0;
    });
}
int at_tree_resize_buffers(at_tree *t)
{
    return guarded([&] { treeOf(t).resizeBuffers(); });
}
int at_tree_buffer(at_tree *t, int id, const void **d, size_t *s, size_t *c)
{
    return guarded([&] {
        needOutputs(d, s, c);
        treeOf(t).buffer(id, d, s, c);
    });
}
int at_tree_scalars(at_tree *t, at_scalars *o)
{
    return guarded([&] {
        needOutputs(o);
        treeOf(t).scalars(o);
    });
}
int at_tree_device_counters(at_tree *t, const int32_t **l, const int32_t **nn, const uint32_t **a)
{
    return guarded([&] {
        needOutputs(l, nn, a);
        at::DevState *s = treeOf(t).deviceState();
        *l = &s->numLeaves;
        *nn = &s->numNodes;
        *a = &s->numActiveParticles;
    });
}
int at_tree_update_passes(at_tree *t, uint32_t *passes)
{
    return guarded([&] {
        needOutputs(passes);
        at::TreeBase &tree = treeOf(t);
        at::DevState h{};
        AT_HIP(hipMemcpyAsync(&h, tree.deviceState(), sizeof(h), hipMemcpyDeviceToHost, tree.stream()));
        at::syncStream(tree.stream());
        *passes = h.passes;
    });
}
int at_tree_max_achieved_depth(at_tree *t, uint32_t *d)
{
    return guarded([&] {
        needOutputs(d);
        *d = treeOf(t).maxAchievedDepth();
    });
}
int at_tree_leaf_count_violations(at_tree *t, uint32_t bucket, uint64_t *overfull, uint64_t *mergeable)
{
    return guarded([&] {
        needOutputs(overfull, mergeable);
        at::TreeBase &tree = treeOf(t);
        const void *K = nullptr, *P = nullptr;
        std::size_t nk = 0, np = 0, cap = 0;
        tree.buffer(AT_BUF_CORNERSTONE, &K, &nk, &cap);
        tree.buffer(AT_BUF_PARTICLE_KEYS, &P, &np, &cap);
        at_scalars sc{};
        tree.scalars(&sc);
        const hipStream_t s = tree.stream();
        at::DevBuf<unsigned long long> counts;
        counts.ensure(2, 0, s);
        AT_HIP(hipMemsetAsync(counts.p, 0, 2 * sizeof(unsigned long long), s));
        at::leafCountViolations(s, static_cast<const at::Key *>(K), nk ? nk - 1 : 0, static_cast<const at::Key *>(P), np, bucket,
                            sc.tree_type == AT_TREE_BINARY ? 1 : 3, counts.p);
        unsigned long long h[2] = {0, 0};
        AT_HIP(hipMemcpyAsync(h, counts.p, sizeof(h), hipMemcpyDeviceToHost, s));
        at::syncStream(s);
        *overfull = h[0];
        *mergeable = h[1];
    });
}
int at_tree_print(at_tree *t, char *buf, size_t len, size_t *needed)
{
    return guarded([&] {
        const std::string text = at::printTree(treeOf(t));
        if (needed) *needed = text.size() + 1;
        if (buf && len)
        {
            const std::size_t k = std::min(len - 1, text.size());
            std::memcpy(buf, text.data(), k);
            buf[k] = '\0';
        }
    });
}
const char *at_last_error(void) { return g_error.c_str(); }

int at_device_count(int *count) { return hipStatus(hipGetDeviceCount(count)); }
int at_device_malloc(void **p, size_t bytes) { return hipStatus(hipMalloc(p, bytes ? bytes : 1)); }
int at_device_free(void *p) { return hipStatus(hipFree(p)); }
int at_memcpy(void *dst, const void *src, size_t bytes, int kind)
{
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                                      : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    return hipStatus(hipMemcpy(dst, src, bytes, k));
}
int at_memcpy_async(void *dst, const void *src, size_t bytes, int kind, void *stream)
{
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                                      : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    return hipStatus(hipMemcpyAsync(dst, src, bytes, k, hipStream_t(stream)));
}
int at_stream_create(void **s)
{
    hipStream_t h = nullptr;
    const int r = hipStatus(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
    *s = h;
    return r;
}
int at_stream_destroy(void *s)
{
    // blocks still pending on the stream become reusable before it goes
    const hipError_t e = at::syncStreamStatus(hipStream_t(s));
    if (e != hipSuccess) return hipStatus(e);
    return hipStatus(hipStreamDestroy(hipStream_t(s)));
}
int at_stream_synchronize(void *s) { return hipStatus(at::syncStreamStatus(hipStream_t(s))); }
int at_device_synchronize(void) { return hipStatus(hipDeviceSynchronize()); }
int at_event_create(void **e)
{
    hipEvent_t h = nullptr;
    const int r = hipStatus(hipEventCreate(&h));
    *e = h;
    return r;
}
int at_event_destroy(void *e) { return hipStatus(hipEventDestroy(hipEvent_t(e))); }
int at_event_record(void *e, void *s) { return hipStatus(hipEventRecord(hipEvent_t(e), hipStream_t(s))); }
int at_event_elapsed_ms(void *a, void *b, float *ms)
{
    return hipStatus(hipEventElapsedTime(ms, hipEvent_t(a), hipEvent_t(b)));
}
int at_graph_capture_update(at_tree *t, void *stream, void **exec)
{
    return guarded([&] {
        needOutputs(exec);
        hipStream_t s = treeOf(t).stream();
        if (!s) throw at::Error(AT_EINVAL, "graph capture needs a tree created on a non-default stream");
        if (stream && hipStream_t(stream) != s)
            throw at::Error(AT_EINVAL, "capture stream must be the tree's configured stream");
        hipGraph_t g = nullptr;
        AT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeGlobal));
        try
        {
            treeOf(t).update();
        }
        catch (...)
        {
            hipGraph_t dummy = nullptr;
            (void)hipStreamEndCapture(s, &dummy);
            if (dummy) (void)hipGraphDestroy(dummy);
            throw;
        }
        AT_HIP(hipStreamEndCapture(s, &g));
        hipGraphExec_t e = nullptr;
        AT_HIP(hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        *exec = e;
    });
}
int at_graph_launch(void *exec, void *stream)
{
    return hipStatus(hipGraphLaunch(hipGraphExec_t(exec), hipStream_t(stream)));
}
int at_graph_destroy(void *exec) { return hipStatus(hipGraphExecDestroy(hipGraphExec_t(exec))); }

int at_tune_set(const char *key, int value)
{
    return guarded([&] {
        const std::string k = key ? key : "";
        if (k == "sort_items") at::setSortItemsPerThread(value);
        else if (k == "sort_ranges") at::setSortTargetRanges(value);
        else if (k == "sort_threads") at::setSortThreads(value);
        else if (k == "sort_digit_array") at::setSortDigitArray(value);
        else if (k == "sort_prefetch") at::setSortPrefetch(value);
        else if (k == "sort_msd") at::setSortMsd(value);
        else if (k == "sort_digit_bits") at::setSortDigitBits(value);
        else if (k == "sort_msd_min_keys") at::setSortMsdMinKeys(value < 0 ? 0 : std::size_t(value));
        else throw at::Error(AT_EINVAL, "unknown tuning parameter: " + k);
    });
}
int at_sort_pairs(void *keys, void *vals, size_t n, int slot_values, void *stream)
{
    return guarded([&] {
        if (n && (!keys || !vals)) throw at::Error(AT_EINVAL, "keys and values must be device pointers");
        at::SortWorkspace ws;
        at::sortPairs(hipStream_t(stream), static_cast<at::Key *>(keys), static_cast<std::uint32_t *>(vals), n, ws,
                      slot_values != 0);
        at::syncStream(hipStream_t(stream)); // before the workspace is released
    });
}
int at_sort_pairs_noalloc(void *keys, void *vals, size_t n, int slot_values, void *stream)
{
    return guarded([&] {
        if (n && (!keys || !vals)) throw at::Error(AT_EINVAL, "keys and values must be device pointers");
        at::sortPairsNoAlloc(hipStream_t(stream), static_cast<at::Key *>(keys), static_cast<std::uint32_t *>(vals), n,
                             slot_values != 0);
        AT_HIP(hipGetLastError());
    });
}
int at_tune_auto(size_t n, void *stream, char *report, size_t report_len)
{
    return guarded([&] {
        const std::string r = at::autotuneSort(n, hipStream_t(stream));
        if (report && report_len)
        {
            const std::size_t m = r.size() < report_len - 1 ? r.size() : report_len - 1;
            std::memcpy(report, r.data(), m);
            report[m] = 0;
        }
    });
}
int at_tune_get(const char *key, int *value)
{
    return guarded([&] {
        needOutputs(value);
        const std::string k = key ? key : "";
        if (k == "sort_items") *value = at::sortItemsPerThread();
        else if (k == "sort_ranges") *value = at::sortTargetRanges();
        else if (k == "sort_threads") *value = at::sortThreads();
        else if (k == "sort_digit_array") *value = at::sortDigitArray();
        else if (k == "sort_prefetch") *value = at::sortPrefetch();
        else if (k == "sort_msd") *value = at::sortMsd();
        else if (k == "sort_digit_bits") *value = at::sortDigitBits();
        else if (k == "sort_msd_min_keys") *value = int(at::sortMsdMinKeys());
        else throw at::Error(AT_EINVAL, "unknown tuning parameter: " + k);
    });
}
}
