// server.hip -- a request loop that keeps trees resident in HBM and serves
// mixed build / update / release requests with continuous batching.
//
//   admission  build requests wait in arrival order until their estimated
//              device bytes fit the HBM budget; a release returns its bytes
//              and admits the waiting builds that now fit.
//   placement  each admitted tree is pinned to one worker (a host thread
//              owning one HIP stream), the worker holding the fewest
//              resident particles (then the fewest trees); all requests of
//              a tree run in submission order on that worker.
//   rounds     a worker takes its queued requests in rounds: in queue
//              order, at most one per tree and none behind a request of its
//              tree left queued (up to max_batch). The round's builds run
//              first, then its updates (new positions copied in), launched
//              on the stream before one synchronisation; then each tree's
//              resize protocol (doBuffersNeedResize / resizeBuffers /
//              update) completes it, and the round's releases free their
//              trees. Workers run concurrently, so rounds of different
//              workers overlap on the device.
//   groups     octrees of one configuration (precision, split criterion,
//              leaf size, key kind, box, limits) of up to 4M particles in a
//              round share their launches (GroupRunner: one segmented sort,
//              multi-tree pass kernels; builds with device-driven passes),
//              in groups of up to 16M particles; the rest run tree by tree.
//   graphs     with update_graphs set, a tree's update() is captured
//              (thread-local capture on its worker's stream) at its first
//              update and again at the first update after a resize, one
//              capture at a time in the process, and replayed as one
//              launch. While any stream is being captured, HIP fails
//              operations on the legacy default stream in every thread of
//              the process (hipErrorStreamCaptureImplicit), so captures
//              run only inside requests and only when the caller asks.
//   release    a tree with nothing queued or running is released at once:
//              the request completes at submission and the worker frees
//              the tree afterwards (its bytes return to the budget then);
//              otherwise the release runs after the tree's requests.
// Every request runs the same tree operations it would run alone, on a tree
// that only its own requests touch, so results do not depend on the mix.
#include "batch.hpp"
#include "engine.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>

namespace at
{
namespace
{
using Clock = std::chrono::steady_clock;

enum Kind { kBuild, kUpdate, kRelease };
enum State { kQueued, kRunning, kDone };

struct Request
{
    std::uint64_t id = 0, tree = 0;
    Kind kind = kBuild;
    bool internal = false; // the deferred destruction of a tree released at submission
    // build
    int realBytes = 8, type = AT_TREE_OCTREE;
    const void *x = nullptr, *y = nullptr, *z = nullptr;
    std::size_t n = 0;
    at_config cfg{};
    // update: new positions copied into the tree's coordinates (may be null)
    const void *nx = nullptr, *ny = nullptr, *nz = nullptr;
    State state = kQueued;
    int status = AT_OK;
    std::string error;
    Clock::time_point submit, start, done;
};

// An instantiated update graph, destroyed with its tree.
struct GraphExec
{
    hipGraphExec_t g = nullptr;
    GraphExec() = default;
    GraphExec(const GraphExec &) = delete;
    GraphExec &operator=(const GraphExec &) = delete;
    GraphExec(GraphExec &&o) noexcept : g(o.g) { o.g = nullptr; }
    GraphExec &operator=(GraphExec &&o) noexcept
    {
        std::swap(g, o.g);
        return *this;
    }
    ~GraphExec() { reset(); }
    void reset()
    {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
};

// Records tree.update() on stream s; null when the capture is refused.
// Captures and instantiations of all servers run one at a time.
std::mutex g_captureMu;
hipGraphExec_t captureUpdate(TreeBase &tree, hipStream_t s)
{
    std::lock_guard<std::mutex> lock(g_captureMu);
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
    bool ok = true;
    try
    {
        tree.update();
    }
    catch (...)
    {
        ok = false;
    }
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(s, &g) != hipSuccess) ok = false;
    hipGraphExec_t e = nullptr;
    if (ok && g && hipGraphInstantiate(&e, g, nullptr, nullptr, 0) != hipSuccess) e = nullptr;
    if (g) (void)hipGraphDestroy(g);
    return ok ? e : nullptr;
}

struct Resident
{
    std::unique_ptr<TreeBase> tree;
    GraphExec graph; // touched only by the tree's worker
    int worker = -1;
    bool admitted = false;
    bool releasing = false; // a release is queued: no further requests
    int outstanding = 0;    // its requests queued on or running at its worker
    bool failed = false;    // the build failed: no tree, bytes returned
    std::size_t bytes = 0, n = 0;
    int realBytes = 8, type = AT_TREE_OCTREE;
    at_config cfg{};
    const void *x = nullptr, *y = nullptr, *z = nullptr;
    std::deque<Request *> held; // requests submitted before the build was admitted
};

// Trees whose requests may share launches: same precision, type and
// configuration (the stream aside).
struct ClassKey
{
    int realBytes, type, split, sfc, maxDepth, iterations;
    unsigned bucket;
    std::array<std::uint8_t, 3> views;
    std::array<double, 6> box;
    explicit ClassKey(int rb, int t, const at_config &c)
        : realBytes(rb), type(t), split(c.split_criterion), sfc(c.sfc_kind), maxDepth(c.max_depth),
          iterations(c.update_iterations), bucket(c.max_particles_per_leaf),
          views{c.views[0], c.views[1], c.views[2]},
          box{c.box.xmin, c.box.xmax, c.box.ymin, c.box.ymax, c.box.zmin, c.box.zmax}
    {
    }
    bool operator<(const ClassKey &o) const
    {
        return std::tie(realBytes, type, split, sfc, maxDepth, iterations, bucket, views, box) <
               std::tie(o.realBytes, o.type, o.split, o.sfc, o.maxDepth, o.iterations, o.bucket, o.views, o.box);
    }
};

std::size_t estimateBytes(std::size_t n, int realBytes, const at_config &c)
{
    // P + Perm + sort scratch (keys, values, digit bytes) + leaves (~n/8
    // leaves x two halves of K, N and the scan) + views when configured
    const bool views = c.views[0] || c.views[1] || c.views[2];
    (void)realBytes;
    return n * (8 + 4 + 13) + (n / 8 + 64) * 48 + (views ? n * 8 + (n / 8) * 26 : 0) + (std::size_t(1) << 20);
}
} // namespace

class Server
{
public:
    explicit Server(const at_server_config &c)
        : budget_(c.hbm_budget_bytes), maxBatch_(c.max_batch > 0 ? c.max_batch : 16), graphs_(c.update_graphs != 0),
          t0_(Clock::now())
    {
        const int w = c.workers > 0 ? c.workers : 4;
        if (w > 32) throw Error(AT_EINVAL, "at most 32 server workers");
        workers_.resize(std::size_t(w));
        for (auto &wk : workers_) AT_HIP(hipStreamCreateWithFlags(&wk.stream, hipStreamNonBlocking));
        for (int i = 0; i < w; ++i) threads_.emplace_back([this, i] { loop(i); });
    }

    // Requests already on a worker run to completion; builds still waiting
    // for budget, and the requests held behind them, complete with
    // AT_ELOGIC. Returns once no thread is inside wait().
    ~Server()
    {
        {
            std::lock_guard<std::mutex> g(m_);
            stopping_ = true;
            const auto now = Clock::now();
            auto abandon = [&](Request *r) {
                r->status = AT_ELOGIC;
                r->error = "server destroyed before the request was admitted";
                r->state = kDone;
                r->start = r->done = now;
            };
            for (Request *r : pendingBuilds_) abandon(r);
            pendingBuilds_.clear();
            for (auto &kv : trees_)
            {
                for (Request *r : kv.second.held) abandon(r);
                kv.second.held.clear();
            }
        }
        cv_.notify_all();
        doneCv_.notify_all();
        for (auto &t : threads_) t.join();
        {
            std::unique_lock<std::mutex> g(m_);
            doneCv_.wait(g, [&] { return waiters_ == 0; });
        }
        trees_.clear();
        for (auto &wk : workers_)
        {
            wk.groups.reset();
            (void)syncStreamStatus(wk.stream); // blocks pending on it become reusable
            (void)hipStreamDestroy(wk.stream);
        }
        for (auto &kv : requests_) delete kv.second;
    }

    std::uint64_t submitBuild(int realBytes, int type, const void *x, const void *y, const void *z, std::size_t n,
                              const at_config &cfg, std::uint64_t *treeId)
    {
        if (realBytes != 4 && realBytes != 8) throw Error(AT_EINVAL, "real_bytes must be 4 or 8");
        if (n > 0 && (!x || !y || !z))
            throw Error(AT_EINVAL, "adaptive_octree: coordinate device pointers must not be null");
        const std::size_t bytes = estimateBytes(n, realBytes, cfg);
        std::lock_guard<std::mutex> g(m_);
        if (bytes > budget_) throw Error(AT_ENOMEM, "tree exceeds the server's HBM budget");
        Request *r = newRequest(kBuild);
        r->tree = ++lastTree_;
        r->realBytes = realBytes;
        r->type = type;
        r->x = x, r->y = y, r->z = z;
        r->n = n;
        r->cfg = cfg;
        Resident &t = trees_[r->tree];
        t.n = n;
        t.realBytes = realBytes;
        t.type = type;
        t.cfg = cfg;
        t.x = x, t.y = y, t.z = z;
        t.bytes = bytes;
        pendingBuilds_.push_back(r);
        admit();
        *treeId = r->tree;
        return r->id;
    }

    std::uint64_t submitUpdate(std::uint64_t tree, const void *nx, const void *ny, const void *nz)
    {
        std::lock_guard<std::mutex> g(m_);
        Resident &t = liveResident(tree);
        Request *r = newRequest(kUpdate);
        r->tree = tree;
        r->nx = nx, r->ny = ny, r->nz = nz;
        route(t, r);
        return r->id;
    }

    std::uint64_t submitRelease(std::uint64_t tree)
    {
        std::lock_guard<std::mutex> g(m_);
        Resident &t = liveResident(tree);
        t.releasing = true;
        Request *r = newRequest(kRelease);
        r->tree = tree;
        if (t.admitted && t.outstanding == 0)
        {
            // nothing of the tree is queued or running: done now, freed by its worker
            r->state = kDone;
            r->start = r->done = r->submit;
            ++completed_;
            Request *d = newRequest(kRelease);
            d->tree = tree;
            d->internal = true;
            enqueue(t.worker, d, 0);
            doneCv_.notify_all();
            return r->id;
        }
        route(t, r);
        return r->id;
    }

    int wait(std::uint64_t id)
    {
        std::unique_lock<std::mutex> g(m_);
        Request *r = requestOf(id);
        ++waiters_;
        doneCv_.wait(g, [&] { return r->state == kDone; });
        const int status = r->status;
        if (status != AT_OK) lastError() = r->error;
        if (--waiters_ == 0 && stopping_) doneCv_.notify_all();
        return status;
    }

    bool poll(std::uint64_t id)
    {
        std::lock_guard<std::mutex> g(m_);
        return requestOf(id)->state == kDone;
    }

    void times(std::uint64_t id, double *submit, double *start, double *done)
    {
        std::lock_guard<std::mutex> g(m_);
        Request *r = requestOf(id);
        auto ms = [&](Clock::time_point t) { return std::chrono::duration<double, std::milli>(t - t0_).count(); };
        if (submit) *submit = ms(r->submit);
        if (start) *start = r->state == kQueued ? -1.0 : ms(r->start);
        if (done) *done = r->state == kDone ? ms(r->done) : -1.0;
    }

    // The tree must have no request in flight (wait for the last one first).
    TreeBase &tree(std::uint64_t id)
    {
        std::lock_guard<std::mutex> g(m_);
        Resident &t = liveResident(id);
        if (t.failed) throw Error(AT_ELOGIC, "tree failed to build");
        if (!t.tree) throw Error(AT_ELOGIC, "tree is not built yet");
        return *t.tree;
    }

    void stats(at_server_stats *o)
    {
        std::lock_guard<std::mutex> g(m_);
        o->bytes_in_use = used_;
        o->bytes_peak = peak_;
        o->hbm_budget = budget_;
        o->resident_trees = trees_.size();
        o->waiting_builds = pendingBuilds_.size();
        o->completed = completed_;
        o->batched_updates = batchedUpdates_;
        o->update_batches = updateBatches_;
        o->grouped_builds = groupedBuilds_;
        o->grouped_updates = groupedUpdates_;
    }

private:
    struct Worker
    {
        hipStream_t stream = nullptr;
        std::deque<Request *> queue;
        std::size_t work = 0;     // queued particles
        std::size_t resident = 0; // particles of the trees pinned here
        std::size_t trees = 0;
        std::unique_ptr<GroupRunner> groups; // made and used by the worker's thread
    };
    static constexpr std::size_t kRoundScan = 4096; // queued requests looked at per round

    Request *newRequest(Kind k)
    {
        Request *r = new Request;
        r->id = ++lastRequest_;
        r->kind = k;
        r->submit = Clock::now();
        requests_[r->id] = r;
        return r;
    }
    Request *requestOf(std::uint64_t id)
    {
        auto it = requests_.find(id);
        if (it == requests_.end()) throw Error(AT_EINVAL, "unknown request id");
        return it->second;
    }
    // A tree that takes requests: known, and no release queued for it.
    Resident &liveResident(std::uint64_t id)
    {
        auto it = trees_.find(id);
        if (it == trees_.end() || it->second.releasing) throw Error(AT_EINVAL, "unknown or released tree id");
        return it->second;
    }

    // Requests of an admitted tree go to its worker; otherwise they wait
    // behind the tree's build.
    void route(Resident &t, Request *r)
    {
        if (!t.admitted)
        {
            t.held.push_back(r);
            return;
        }
        enqueue(t.worker, r, t.n);
    }
    void enqueue(int w, Request *r, std::size_t n)
    {
        workers_[std::size_t(w)].queue.push_back(r);
        workers_[std::size_t(w)].work += n;
        ++trees_.at(r->tree).outstanding;
        cv_.notify_all();
    }

    // Admits waiting builds in arrival order while their bytes fit.
    void admit()
    {
        while (!pendingBuilds_.empty())
        {
            Request *r = pendingBuilds_.front();
            Resident &t = trees_.at(r->tree);
            if (used_ + t.bytes > budget_) break;
            pendingBuilds_.pop_front();
            used_ += t.bytes;
            peak_ = std::max(peak_, used_);
            int best = 0;
            for (int w = 1; w < int(workers_.size()); ++w)
            {
                const Worker &a = workers_[std::size_t(w)], &b = workers_[std::size_t(best)];
                if (a.resident < b.resident || (a.resident == b.resident && a.trees < b.trees)) best = w;
            }
            t.worker = best;
            workers_[std::size_t(best)].resident += t.n;
            ++workers_[std::size_t(best)].trees;
            t.admitted = true;
            enqueue(best, r, t.n);
            while (!t.held.empty())
            {
                enqueue(best, t.held.front(), t.n);
                t.held.pop_front();
            }
        }
    }

    // The next requests to run together: in queue order, at most one per
    // tree, and none behind a request of its tree that stays queued (each
    // tree's requests keep their order); up to max_batch of them.
    void takeRound(Worker &wk, std::vector<Request *> &round)
    {
        std::vector<std::uint64_t> seen; // trees with a request taken or passed over
        std::size_t scanned = 0;
        for (auto it = wk.queue.begin();
             it != wk.queue.end() && int(round.size()) < maxBatch_ && scanned < kRoundScan; ++scanned)
        {
            Request *q = *it;
            if (std::find(seen.begin(), seen.end(), q->tree) != seen.end())
            {
                ++it;
                continue;
            }
            seen.push_back(q->tree);
            round.push_back(q);
            it = wk.queue.erase(it);
        }
    }

    // Members of one class in groups of at most GroupRunner::maxParticles()
    // particles; a group of one runs on its own.
    template <class Size>
    static void formGroups(const std::map<ClassKey, std::vector<std::size_t>> &byClass, Size sizeOf,
                           std::vector<std::vector<std::size_t>> &groups, std::vector<std::size_t> &alone)
    {
        for (const auto &kv : byClass)
        {
            std::vector<std::size_t> g;
            std::size_t total = 0;
            auto flush = [&] {
                if (g.size() > 1) groups.push_back(g);
                else if (g.size() == 1) alone.push_back(g[0]);
                g.clear();
                total = 0;
            };
            for (std::size_t i : kv.second)
            {
                if (total + sizeOf(i) > GroupRunner::maxParticles()) flush();
                g.push_back(i);
                total += sizeOf(i);
            }
            flush();
        }
    }

    void loop(int w)
    {
        Worker &wk = workers_[std::size_t(w)];
        for (;;)
        {
            std::vector<Request *> batch;
            std::vector<Resident *> res;
            std::size_t updates = 0;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return stopping_ || !wk.queue.empty(); });
                if (wk.queue.empty()) return; // stopping and drained
                takeRound(wk, batch);
                const auto now = Clock::now();
                for (Request *r : batch)
                {
                    r->state = kRunning;
                    r->start = now;
                    // every queued request's tree is resident: its release
                    // is queued behind it and refuses later submissions
                    Resident &t = trees_.at(r->tree);
                    res.push_back(&t);
                    wk.work -= std::min(wk.work, r->internal ? std::size_t(0) : t.n);
                    updates += r->kind == kUpdate ? 1 : 0;
                }
                if (updates > 1)
                {
                    ++updateBatches_;
                    batchedUpdates_ += updates;
                }
            }
            const std::size_t R = batch.size();
            std::vector<int> status(R, AT_OK);
            std::vector<std::string> errors(R);
            auto fail = [&](std::size_t i, int code, const char *what) {
                status[i] = code;
                errors[i] = what;
            };
            auto guarded = [&](std::size_t i, auto &&f) {
                try
                {
                    f();
                    return true;
                }
                catch (const Error &e) { fail(i, e.code, e.what()); }
                catch (const std::exception &e) { fail(i, AT_ENOMEM, e.what()); }
                return false;
            };
            if (!wk.groups) wk.groups.reset(new GroupRunner(wk.stream));
            GroupRunner &runner = *wk.groups;

            // ---- builds: every tree made, then groups of one class built together
            std::vector<std::unique_ptr<TreeBase>> built(R);
            std::map<ClassKey, std::vector<std::size_t>> byClass;
            std::vector<std::size_t> alone;
            for (std::size_t i = 0; i < R; ++i)
            {
                Request *r = batch[i];
                if (r->kind != kBuild) continue;
                const bool made = guarded(i, [&] {
                    at_config c = r->cfg;
                    c.stream = wk.stream;
                    built[i].reset(createTree(r->realBytes, r->type, r->x, r->y, r->z, r->n, c));
                });
                if (!made) continue;
                if (!graphs_ && GroupRunner::groupable(r->type, r->cfg, r->n))
                    byClass[ClassKey(r->realBytes, r->type, r->cfg)].push_back(i);
                else
                    alone.push_back(i);
            }
            std::vector<std::vector<std::size_t>> groups;
            formGroups(byClass, [&](std::size_t i) { return batch[i]->n; }, groups, alone);
            std::size_t groupedBuilds = 0, groupedUpdates = 0;
            for (const auto &g : groups)
            {
                groupedBuilds += g.size();
                std::vector<TreeBase *> members;
                for (std::size_t i : g) members.push_back(built[i].get());
                bool ok = false;
                std::string what;
                int code = AT_OK;
                try
                {
                    runner.build(members);
                    ok = true;
                }
                catch (const Error &e) { code = e.code, what = e.what(); }
                catch (const std::exception &e) { code = AT_ENOMEM, what = e.what(); }
                if (!ok)
                    for (std::size_t i : g) fail(i, code, what.c_str());
            }
            for (std::size_t i : alone)
                (void)guarded(i, [&] {
                    built[i]->build();
                    (void)built[i]->needsResize();
                });

            // ---- updates: new positions copied in, groups and single trees
            // launched, one synchronisation, then each tree's resize protocol
            std::vector<bool> launched(R, false);
            std::vector<char> resize(R, 0);
            byClass.clear();
            alone.clear();
            for (std::size_t i = 0; i < R; ++i)
            {
                Request *r = batch[i];
                if (r->kind != kUpdate) continue;
                const bool copied = guarded(i, [&] {
                    if (!res[i]->tree) throw Error(AT_ELOGIC, "tree failed to build");
                    const std::size_t bytes = res[i]->n * std::size_t(res[i]->realBytes);
                    if (r->nx && bytes)
                    {
                        AT_HIP(hipMemcpyAsync(const_cast<void *>(res[i]->x), r->nx, bytes, hipMemcpyDeviceToDevice,
                                              wk.stream));
                        AT_HIP(hipMemcpyAsync(const_cast<void *>(res[i]->y), r->ny, bytes, hipMemcpyDeviceToDevice,
                                              wk.stream));
                        AT_HIP(hipMemcpyAsync(const_cast<void *>(res[i]->z), r->nz, bytes, hipMemcpyDeviceToDevice,
                                              wk.stream));
                    }
                });
                if (!copied) continue;
                if (!graphs_ && GroupRunner::groupable(res[i]->type, res[i]->cfg, res[i]->n))
                    byClass[ClassKey(res[i]->realBytes, res[i]->type, res[i]->cfg)].push_back(i);
                else
                    alone.push_back(i);
            }
            groups.clear();
            formGroups(byClass, [&](std::size_t i) { return res[i]->n; }, groups, alone);
            std::vector<std::vector<std::size_t>> launchedGroups;
            for (const auto &g : groups)
            {
                std::vector<TreeBase *> members;
                for (std::size_t i : g) members.push_back(res[i]->tree.get());
                std::string what;
                int code = AT_OK;
                try
                {
                    runner.launchUpdate(members);
                    groupedUpdates += g.size();
                    launchedGroups.push_back(g);
                    for (std::size_t i : g) launched[i] = true;
                    continue;
                }
                catch (const Error &e) { code = e.code, what = e.what(); }
                catch (const std::exception &e) { code = AT_ENOMEM, what = e.what(); }
                for (std::size_t i : g) fail(i, code, what.c_str());
            }
            for (std::size_t i : alone)
                launched[i] = guarded(i, [&] {
                    TreeBase *t = res[i]->tree.get();
                    if (graphs_ && !res[i]->graph.g) res[i]->graph.g = captureUpdate(*t, wk.stream);
                    if (res[i]->graph.g) AT_HIP(hipGraphLaunch(res[i]->graph.g, wk.stream));
                    else t->update();
                    resize[i] = 1; // asks needsResize after the synchronisation
                });
            bool readBack = true;
            try
            {
                runner.readBack();
            }
            catch (const std::exception &)
            {
                readBack = false;
            }
            // an asynchronous failure belongs to any launched request
            const hipError_t synced = syncStreamStatus(wk.stream);
            if (synced != hipSuccess || !readBack)
            {
                runner.discardUpdates();
                const std::string what = synced != hipSuccess
                                             ? std::string("update batch failed: ") + hipGetErrorString(synced)
                                             : std::string("update batch failed: state read-back");
                for (std::size_t i = 0; i < R; ++i)
                    if (launched[i])
                    {
                        status[i] = synced != hipSuccess ? AT_EHIP : AT_ENOMEM;
                        errors[i] = what;
                        launched[i] = false;
                    }
            }
            else if (!launchedGroups.empty())
            {
                std::vector<std::vector<char>> flags;
                runner.finishUpdates(flags);
                for (std::size_t g = 0; g < launchedGroups.size(); ++g)
                    for (std::size_t k = 0; k < launchedGroups[g].size(); ++k)
                        resize[launchedGroups[g][k]] = flags[g][k];
            }
            for (std::size_t i = 0; i < R; ++i)
            {
                if (!launched[i] || !resize[i]) continue;
                (void)guarded(i, [&] {
                    TreeBase *t = res[i]->tree.get();
                    int attempts = 0;
                    while (t->needsResize())
                    {
                        if (++attempts > 64) throw Error(AT_ELENGTH, "update did not fit after 64 resizes");
                        res[i]->graph.reset(); // the resize moves the captured buffers: captured anew next time
                        t->resizeBuffers();
                        t->update();
                    }
                });
            }

            // ---- completion: trees kept, failed builds and releases returned
            std::vector<std::unique_ptr<TreeBase>> dropped;
            {
                std::lock_guard<std::mutex> g(m_);
                const auto now = Clock::now();
                for (std::size_t i = 0; i < R; ++i)
                {
                    Request *r = batch[i];
                    if (r->kind != kRelease) --res[i]->outstanding;
                    if (r->kind == kBuild && status[i] == AT_OK) res[i]->tree = std::move(built[i]);
                    else if (r->kind == kBuild)
                    {
                        // no tree is kept; its bytes go back to the budget
                        dropped.push_back(std::move(built[i]));
                        res[i]->failed = true;
                        used_ -= std::min(used_, res[i]->bytes);
                        res[i]->bytes = 0;
                        wk.resident -= std::min(wk.resident, res[i]->n);
                        wk.trees -= std::min<std::size_t>(wk.trees, 1);
                    }
                    else if (r->kind == kRelease)
                    {
                        auto it = trees_.find(r->tree);
                        used_ -= std::min(used_, it->second.bytes);
                        if (!it->second.failed)
                        {
                            wk.resident -= std::min(wk.resident, it->second.n);
                            wk.trees -= std::min<std::size_t>(wk.trees, 1);
                        }
                        it->second.graph.reset();
                        dropped.push_back(std::move(it->second.tree));
                        trees_.erase(it);
                    }
                }
                groupedBuilds_ += groupedBuilds;
                groupedUpdates_ += groupedUpdates;
                for (std::size_t i = 0; i < R; ++i)
                {
                    batch[i]->status = status[i];
                    batch[i]->error = errors[i];
                    batch[i]->state = kDone;
                    batch[i]->done = now;
                    if (!batch[i]->internal) ++completed_;
                }
                admit();
            }
            dropped.clear(); // free device memory outside the lock
            doneCv_.notify_all();
        }
    }

    std::size_t budget_ = 0, used_ = 0, peak_ = 0;
    int maxBatch_ = 16;
    bool graphs_ = true;
    Clock::time_point t0_;
    std::mutex m_;
    std::condition_variable cv_, doneCv_;
    bool stopping_ = false;
    int waiters_ = 0; // threads inside wait()
    std::vector<Worker> workers_;
    std::vector<std::thread> threads_;
    std::map<std::uint64_t, Resident> trees_;
    std::map<std::uint64_t, Request *> requests_;
    std::deque<Request *> pendingBuilds_;
    std::uint64_t lastTree_ = 0, lastRequest_ = 0;
    std::uint64_t completed_ = 0, batchedUpdates_ = 0, updateBatches_ = 0;
    std::uint64_t groupedBuilds_ = 0, groupedUpdates_ = 0;
};
} // namespace at

// ---------------------------------------------------------------------------
// C ABI of the server
// ---------------------------------------------------------------------------
struct at_server
{
    std::unique_ptr<at::Server> impl;
};

namespace
{
template <class F>
int guardServer(F &&f)
{
    try
    {
        f();
        return AT_OK;
    }
    catch (const at::Error &e)
    {
        at::lastError() = e.what();
        return e.code;
    }
    catch (const std::bad_alloc &e)
    {
        at::lastError() = e.what();
        return AT_ENOMEM;
    }
    catch (const std::exception &e)
    {
        at::lastError() = e.what();
        return AT_EINVAL;
    }
}
at::Server &serverOf(at_server *s)
{
    if (!s || !s->impl) throw at::Error(AT_EINVAL, "null server");
    return *s->impl;
}
} // namespace

extern "C" {
int at_server_create(const at_server_config *config, at_server **out)
{
    return guardServer([&] {
        if (!config || !out) throw at::Error(AT_EINVAL, "null argument");
        auto s = std::make_unique<at_server>();
        s->impl = std::make_unique<at::Server>(*config);
        *out = s.release();
    });
}
int at_server_destroy(at_server *s)
{
    return guardServer([&] { delete s; });
}
int at_server_submit_build(at_server *s, int realBytes, int type, const void *x, const void *y, const void *z,
                           size_t count, const at_config *config, uint64_t *tree, uint64_t *request)
{
    return guardServer([&] {
        if (!config || !tree || !request) throw at::Error(AT_EINVAL, "null argument");
        *request = serverOf(s).submitBuild(realBytes, type, x, y, z, count, *config, tree);
    });
}
int at_server_submit_update(at_server *s, uint64_t tree, const void *nx, const void *ny, const void *nz,
                            uint64_t *request)
{
    return guardServer([&] {
        if (!request) throw at::Error(AT_EINVAL, "null argument");
        if ((nx || ny || nz) && !(nx && ny && nz)) throw at::Error(AT_EINVAL, "give all three new coordinate arrays");
        *request = serverOf(s).submitUpdate(tree, nx, ny, nz);
    });
}
int at_server_submit_release(at_server *s, uint64_t tree, uint64_t *request)
{
    return guardServer([&] {
        if (!request) throw at::Error(AT_EINVAL, "null argument");
        *request = serverOf(s).submitRelease(tree);
    });
}
int at_server_wait(at_server *s, uint64_t request, int *status)
{
    return guardServer([&] {
        const int st = serverOf(s).wait(request);
        if (status) *status = st;
    });
}
int at_server_poll(at_server *s, uint64_t request, int *done)
{
    return guardServer([&] {
        if (!done) throw at::Error(AT_EINVAL, "null argument");
        *done = serverOf(s).poll(request) ? 1 : 0;
    });
}
int at_server_request_times(at_server *s, uint64_t request, double *submit_ms, double *start_ms, double *done_ms)
{
    return guardServer([&] { serverOf(s).times(request, submit_ms, start_ms, done_ms); });
}
int at_server_tree_buffer(at_server *s, uint64_t tree, int id, const void **d, size_t *size, size_t *cap)
{
    return guardServer([&] {
        if (!d || !size || !cap) throw at::Error(AT_EINVAL, "null output pointer");
        serverOf(s).tree(tree).buffer(id, d, size, cap);
    });
}
int at_server_tree_scalars(at_server *s, uint64_t tree, at_scalars *o)
{
    return guardServer([&] {
        if (!o) throw at::Error(AT_EINVAL, "null output pointer");
        serverOf(s).tree(tree).scalars(o);
    });
}
int at_server_get_stats(at_server *s, at_server_stats *o)
{
    return guardServer([&] {
        if (!o) throw at::Error(AT_EINVAL, "null output pointer");
        serverOf(s).stats(o);
    });
}
}
