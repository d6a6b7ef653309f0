// batch.hip -- hundreds of independent octrees built and updated together.
//
// A batch owns T member trees (the same Octree<Real> objects as the single
// tree API, with the same capacity rules) and drives them through every
// stage with one launch for all members:
//   keys + sort   one segmented key generation and radix sort over the
//                 concatenated particles (SegmentedSort): each member's
//                 P/Perm are its segment of the batch arrays;
//   passes        multi-tree kernels (tree_kernels.hip) over a TreeSet: a
//                 table of member TreeDevs plus, per launch shape, the
//                 prefix of per-member block counts, so small and large
//                 trees share one grid with no idle blocks.
// The build runs every member's passes on the device with no host read-back
// in between: each member's buffers are reserved for a leaf count its build
// is not expected to exceed, and before each step of a pass a one-block plan
// kernel turns the members' device states into the next launches' block
// maps (members that converged get none). A member leaves where its own
// build() would stop, the pass that changes nothing; the host reads every
// state once at the end and replays the committed passes' leaf counts into
// the mirrored buffer capacities. A member whose pass would overflow its
// reservation (rejected on the device, its tree unchanged) finishes with
// host-driven lock-step passes. update() has fixed capacity-sized tables and
// is graph-capturable.
// Results are those of each member's own build()/update(): the per-tree
// arithmetic is unchanged, only the launches are shared.
#include "batch.hpp"
#include "tree_impl.hpp"

#include <algorithm>
#include <climits>

namespace at
{
namespace
{
// Launch shapes of one TreeSet table.
enum Shape { kClear = 0, kTiles, kLeaves, kNf, kIndex, kTileFirst, kTilesOfP, kShapes };

// Device + pinned host staging of a TreeDev table and its block maps.
struct SetTables
{
    std::size_t cap = 0;
    TreeDev *dTrees = nullptr, *hTrees = nullptr;
    std::uint32_t *dMaps = nullptr, *hMaps = nullptr;
    hipEvent_t uploaded = nullptr; // after the last copy out of the pinned tables
    ~SetTables() { release(); }
    void release()
    {
        if (dTrees) (void)hipFree(dTrees);
        if (dMaps) (void)hipFree(dMaps);
        if (hTrees) (void)hipHostFree(hTrees);
        if (hMaps) (void)hipHostFree(hMaps);
        if (uploaded) (void)hipEventDestroy(uploaded);
        dTrees = hTrees = nullptr;
        dMaps = hMaps = nullptr;
        uploaded = nullptr;
        cap = 0;
    }
    // The pinned tables may be rewritten once their last copy has run (the
    // device tables are rewritten in stream order, after the kernels that
    // read them).
    void waitWritable()
    {
        if (uploaded) AT_HIP(hipEventSynchronize(uploaded));
    }
    void reserve(std::size_t n)
    {
        if (n <= cap) return;
        release();
        cap = n;
        AT_HIP(hipMalloc(&dTrees, n * sizeof(TreeDev)));
        AT_HIP(hipMalloc(&dMaps, kShapes * (n + 1) * sizeof(std::uint32_t)));
        AT_HIP(hipHostMalloc(&hTrees, n * sizeof(TreeDev)));
        AT_HIP(hipHostMalloc(&hMaps, kShapes * (n + 1) * sizeof(std::uint32_t)));
        AT_HIP(hipEventCreateWithFlags(&uploaded, hipEventDisableTiming));
    }
    // hTrees[0..count) and the per-member block counts are filled by the
    // caller through `blocks`; writes the prefixes and copies both tables.
    template <class F>
    void upload(hipStream_t s, std::size_t count, F blocks, TreeSet out[kShapes])
    {
        for (int sh = 0; sh < kShapes; ++sh)
        {
            std::uint32_t *m = hMaps + sh * (cap + 1);
            m[0] = 0;
            for (std::size_t k = 0; k < count; ++k) m[k + 1] = m[k] + blocks(Shape(sh), k);
            out[sh].trees = dTrees;
            out[sh].start = dMaps + sh * (cap + 1);
            out[sh].count = int(count);
            out[sh].blocks = count ? m[count] : 0;
        }
        if (!count) return;
        AT_HIP(hipMemcpyAsync(dTrees, hTrees, count * sizeof(TreeDev), hipMemcpyHostToDevice, s));
        AT_HIP(hipMemcpyAsync(dMaps, hMaps, kShapes * (cap + 1) * sizeof(std::uint32_t), hipMemcpyHostToDevice, s));
        AT_HIP(hipEventRecord(uploaded, s));
    }
};

template <class Real>
class OctreeBatch final : public BatchBase
{
    using Member = impl::OctreeImpl<Real>;

public:
    OctreeBatch(std::size_t T, const Real *const *x, const Real *const *y, const Real *const *z,
                const std::size_t *counts, const at_config &c, unsigned maxDepth, unsigned iterations)
        : stream_(hipStream_t(c.stream)), maxDepth_(maxDepth), iterations_(iterations)
    {
        counts_.assign(counts, counts + T);
        sort_.plan(counts_, stream_);
        const std::size_t total = sort_.total();
        P_.ensure(total ? total : 1, 0, stream_);
        perm_.ensure(total ? total : 1, 0, stream_);
        AT_HIP(hipMalloc(&st_, (T ? T : 1) * sizeof(DevState)));
        AT_HIP(hipHostMalloc(&sth_, (T ? T : 1) * sizeof(DevState)));
        AT_HIP(hipMemsetAsync(st_, 0, (T ? T : 1) * sizeof(DevState), stream_));
        std::vector<SegCoords<Real>> coords(T);
        trees_.reserve(T);
        for (std::size_t k = 0; k < T; ++k)
        {
            coords[k] = SegCoords<Real>{x[k], y[k], z[k]};
            MemberSlot slot{st_ + k, sth_ + k, P_.p + sort_.segBegin(k), perm_.p + sort_.segBegin(k)};
            trees_.emplace_back(new Member(x[k], y[k], z[k], counts[k], c, maxDepth, iterations, &slot));
        }
        coords_.ensure(T ? T : 1, 0, stream_);
        syncStream(stream_);
        if (T) AT_HIP(hipMemcpy(coords_.p, coords.data(), T * sizeof(SegCoords<Real>), hipMemcpyHostToDevice));
        kp_ = KeygenParams<Real>{};
        kp_.grid = impl::makeGridBox<Real>(c.box);
        kind_ = c.sfc_kind == AT_HILBERT ? KeyKind::Hilbert : KeyKind::Morton;
        params_ = OctParams{c.max_particles_per_leaf, c.split_criterion == AT_NF_COUNT, c.sfc_kind == AT_HILBERT};
        build_.reserve(T ? T : 1);
        aux_.reserve(T ? T : 1);
        upd_.reserve(T ? T : 1);
    }

    ~OctreeBatch() override
    {
        (void)syncStreamStatus(stream_);
        trees_.clear();
        if (st_) (void)hipFree(st_);
        if (sth_) (void)hipHostFree(sth_);
    }

    void build() override
    {
        syncStream(stream_);
        built_ = false;
        const std::size_t T = trees_.size();
        sort_.keygenAndSort<Real>(stream_, kind_, kp_, coords_.p, P_.p, perm_.p, false);
        std::vector<std::size_t> all(T);
        for (std::size_t k = 0; k < T; ++k)
        {
            all[k] = k;
            trees_[k]->buildBegin();
            trees_[k]->reserveDeviceBuild(trees_[k]->deviceBuildCapacity());
        }
        // ---- every pass on the device: tables staged once, block maps planned per step
        TreeSet set[kShapes];
        const std::size_t logStride = maxDepth_ + 1;
        passLog_.ensure(T * logStride + 1, 0, stream_);
        stageWith(build_, all, [&](std::size_t i, TreeDev &t) {
            t.capacity = trees_[all[i]]->deviceBuildLimit();
            t.passLog = passLog_.p + all[i] * logStride;
        }, [&](Shape sh, std::size_t k) { return blocksFor(sh, k, 1); }, set);
        multiBuildSearchIndex(stream_, set[kIndex]);
        multiOctInit(stream_, set[kClear]);
        if (T)
        {
            planStart_.ensure(std::size_t(kPlanShapes) * (T + 1), 0, stream_);
            const TreeDev *dt = build_.dTrees;
            const int count = int(T);
            const TreeSet tiles = plannedSet(dt, count, planStart_.p, kPlanTiles);
            const TreeSet proc = plannedSet(dt, count, planStart_.p, kPlanProcess);
            const TreeSet commit = plannedSet(dt, count, planStart_.p, kPlanCommit);
            const TreeSet fresh = plannedSet(dt, count, planStart_.p, kPlanLeaves);
            const TreeSet nfWork = plannedSet(dt, count, planStart_.p, kPlanNf);
            for (unsigned pass = 0; pass < maxDepth_; ++pass)
            {
                multiPlan(stream_, dt, count, 0, params_.nf, planStart_.p);
                multiOctDecideScan(stream_, tiles, params_);
                multiPlan(stream_, dt, count, 1, params_.nf, planStart_.p);
                multiOctProcess(stream_, proc, params_);
                multiCommit(stream_, commit, false);
                multiOctRefresh(stream_, fresh, nfWork, params_, false);
            }
        }
        // ---- the one read-back: states and the committed passes' leaf counts
        logHost_.resize(T * logStride + 1);
        if (T)
        {
            AT_HIP(hipMemcpyAsync(sth_, st_, T * sizeof(DevState), hipMemcpyDeviceToHost, stream_));
            AT_HIP(hipMemcpyAsync(logHost_.data(), passLog_.p, T * logStride * sizeof(std::uint32_t),
                                  hipMemcpyDeviceToHost, stream_));
        }
        syncStream(stream_);
        std::vector<std::size_t> stalled;
        std::vector<unsigned> decided(T, 0);
        for (std::size_t k = 0; k < T; ++k)
        {
            const DevState &st = sth_[k];
            const unsigned passes = std::min(st.passes, maxDepth_);
            for (unsigned j = 0; j < passes; ++j) trees_[k]->replayBuildPass(logHost_[k * logStride + j]);
            decided[k] = passes;
            if (st.needsResize)
            {
                sth_[k].needsResize = 0;
                sth_[k].converged = 0;
                stalled.push_back(k);
            }
        }
        if (!stalled.empty()) hostPasses(stalled, decided);
        for (auto &t : trees_) t->buildEnd(false);
        for (std::size_t k = 0; k < T; ++k) Member::finishState(sth_[k], trees_[k]->buildLeaves());
        if (T) AT_HIP(hipMemcpyAsync(st_, sth_, T * sizeof(DevState), hipMemcpyHostToDevice, stream_));
        built_ = true;
        refreshUpdateTables();
        for (std::size_t k = 0; k < T; ++k) (void)trees_[k]->applyState(sth_[k]);
        for (auto &t : trees_)
            if (t->configuredViews().any()) t->buildViews(t->configuredViews());
        // completes the buffers' growth copies and the uploads above
        syncStream(stream_);
    }

    // Host-driven lock-step passes for `members` (their device states with
    // cleared flags are in sth_), each until its pass changes nothing or
    // maxDepth decisions; decided[k] = decisions already made.
    void hostPasses(std::vector<std::size_t> active, std::vector<unsigned> &decided)
    {
        AT_HIP(hipMemcpyAsync(st_, sth_, trees_.size() * sizeof(DevState), hipMemcpyHostToDevice, stream_));
        TreeSet set[kShapes];
        for (;;)
        {
            active.erase(std::remove_if(active.begin(), active.end(),
                                        [&](std::size_t k) { return decided[k] >= maxDepth_; }),
                         active.end());
            if (active.empty()) break;
            for (std::size_t k : active) trees_[k]->buildPrePass();
            stage(build_, active, true,
                  [&](Shape sh, std::size_t i) { return blocksFor(sh, active[i], trees_[active[i]]->buildLeaves()); },
                  set);
            multiClearScan(stream_, set[kClear]);
            multiOctDecideScan(stream_, set[kTiles], params_);
            readStates();
            std::vector<std::size_t> next;
            for (std::size_t k : active)
            {
                ++decided[k];
                if (trees_[k]->buildAfterDecide(sth_[k])) next.push_back(k);
            }
            if (next.empty()) break;
            // the process grid covers the old leaves, the refresh the new ones
            TreeSet proc[kShapes];
            stage(build_, next, true,
                  [&](Shape sh, std::size_t i) { return blocksFor(sh, next[i], trees_[next[i]]->buildLeaves()); },
                  proc);
            multiOctProcess(stream_, proc[kLeaves], params_);
            multiCommit(stream_, proc[kClear], false);
            TreeSet fresh[kShapes];
            stage(aux_, next, true,
                  [&](Shape sh, std::size_t i) { return blocksFor(sh, next[i], trees_[next[i]]->buildNewLeaves()); },
                  fresh);
            multiOctRefresh(stream_, fresh[kLeaves], fresh[kNf], params_, false);
            for (std::size_t k : next) trees_[k]->buildAfterProcess();
            active.swap(next);
        }
        readStates();
    }

    void update() override
    {
        if (!built_) throw Error(AT_ELOGIC, "adaptive_octree: call build() before update()");
        sort_.keygenAndSort<Real>(stream_, kind_, kp_, coords_.p, P_.p, perm_.p);
        multiBuildSearchIndex(stream_, updSet_[kIndex]);
        multiReset(stream_, updSet_[kClear]);
        // every member recounts all leaves: by tiles of P for small LeafCount
        // leaves (the members' own update() does the same)
        if (!params_.nf && params_.bucket <= kTileCountBucket)
            multiLeafCountsByTiles(stream_, updSet_[kTileFirst], updSet_[kTilesOfP], updSet_[kLeaves]);
        else
            multiOctRefresh(stream_, updSet_[kLeaves], updSet_[kNf], params_, true);
        for (unsigned i = 0; i < iterations_; ++i)
        {
            multiOctDecideScan(stream_, updSet_[kTiles], params_);
            multiOctProcess(stream_, updSet_[kLeaves], params_);
            multiCommit(stream_, updSet_[kClear], false);
            multiOctRefresh(stream_, updSet_[kLeaves], updSet_[kNf], params_, false);
        }
        for (auto &t : trees_)
            if (t->configuredViews().any()) t->buildViews(t->configuredViews());
        AT_HIP(hipGetLastError());
    }

    std::size_t needsResize(int *flags) override
    {
        if (!built_) return 0;
        readStates();
        std::size_t n = 0;
        for (std::size_t k = 0; k < trees_.size(); ++k)
        {
            const bool r = trees_[k]->applyState(sth_[k]);
            if (flags) flags[k] = r ? 1 : 0;
            n += r ? 1 : 0;
        }
        return n;
    }

    void resizeBuffers() override
    {
        if (!built_) throw Error(AT_ELOGIC, "adaptive_octree: call build() before resizeBuffers()");
        readStates();
        for (std::size_t k = 0; k < trees_.size(); ++k)
            if (trees_[k]->applyState(sth_[k]))
            {
                trees_[k]->growForResize();
                AT_HIP(hipMemsetAsync(&st_[k].needsResize, 0, sizeof(int), stream_));
            }
        syncStream(stream_);
        refreshUpdateTables();
    }

    std::size_t size() const override { return trees_.size(); }
    TreeBase &tree(std::size_t k) override
    {
        if (k >= trees_.size()) throw Error(AT_EINVAL, "batch tree index out of range");
        return *trees_[k];
    }
    hipStream_t stream() const override { return stream_; }

private:
    std::uint32_t blocksFor(Shape sh, std::size_t k, std::size_t leaves) const
    {
        switch (sh)
        {
        case kClear: return clearBlocksFor(trees_[k]->device().statusWords);
        case kTiles: return tileBlocksFor(leaves);
        case kLeaves: return leafBlocksFor(leaves);
        case kIndex: return indexBlocksFor(counts_[k]);
        case kTileFirst: return tileFirstBlocksFor(counts_[k]);
        case kTilesOfP: return tileBlocksOfP(counts_[k]);
        default: return params_.nf ? nfBlocksFor(leaves) : 0u;
        }
    }

    // Stages the TreeDevs of `members` (each adjusted by `adjust(i, t)`) and
    // the block maps; one copy of each table.
    template <class A, class F>
    void stageWith(SetTables &tab, const std::vector<std::size_t> &members, A adjust, F blocks,
                   TreeSet out[kShapes])
    {
        tab.waitWritable();
        for (std::size_t i = 0; i < members.size(); ++i)
        {
            TreeDev t = trees_[members[i]]->device();
            adjust(i, t);
            tab.hTrees[i] = t;
        }
        tab.upload(stream_, members.size(), blocks, out);
    }
    // build: capacity unlimited (host-driven passes grow the buffers)
    template <class F>
    void stage(SetTables &tab, const std::vector<std::size_t> &members, bool building, F blocks, TreeSet out[kShapes])
    {
        stageWith(tab, members, [&](std::size_t, TreeDev &t) {
            if (building) t.capacity = LONG_MAX;
        }, blocks, out);
    }

    void refreshUpdateTables()
    {
        std::vector<std::size_t> all(trees_.size());
        for (std::size_t k = 0; k < all.size(); ++k) all[k] = k;
        stage(upd_, all, false,
              [&](Shape sh, std::size_t k) { return blocksFor(sh, k, std::size_t(trees_[k]->device().capacity)); },
              updSet_);
    }

    void readStates()
    {
        if (!trees_.empty())
            AT_HIP(hipMemcpyAsync(sth_, st_, trees_.size() * sizeof(DevState), hipMemcpyDeviceToHost, stream_));
        syncStream(stream_);
    }

    hipStream_t stream_ = nullptr;
    unsigned maxDepth_ = 21, iterations_ = 2;
    bool built_ = false;
    std::vector<std::size_t> counts_;
    std::vector<std::unique_ptr<Member>> trees_;
    SegmentedSort sort_;
    DevBuf<Key> P_;
    DevBuf<std::uint32_t> perm_;
    DevBuf<SegCoords<Real>> coords_;
    DevState *st_ = nullptr, *sth_ = nullptr;
    KeygenParams<Real> kp_{};
    KeyKind kind_ = KeyKind::Morton;
    OctParams params_{};
    SetTables build_, aux_, upd_;
    TreeSet updSet_[kShapes];
    DevBuf<std::uint32_t> planStart_; // [kPlanShapes][T + 1] block maps planned on the device
    DevBuf<std::uint32_t> passLog_;   // [T][maxDepth + 1] leaf counts of the committed build passes
    std::vector<std::uint32_t> logHost_;
};

template <class Real>
BatchBase *makeBatch(int type, std::size_t T, const void *const *x, const void *const *y, const void *const *z,
                     const std::size_t *counts, const at_config &c)
{
    if (type != AT_TREE_OCTREE) throw Error(AT_EINVAL, "batches hold octrees");
    if (c.max_depth < -1 || c.update_iterations < -1) throw Error(AT_EINVAL, "Tree: use -1 for backend defaults");
    if (c.max_particles_per_leaf == 0) throw Error(AT_EINVAL, "adaptive_octree: particle limit must be positive");
    const unsigned it = c.update_iterations >= 0 ? unsigned(c.update_iterations) : 2u;
    if (it == 0) throw Error(AT_EINVAL, "adaptive_octree: updateIterations must be positive");
    if (T && (!x || !y || !z || !counts)) throw Error(AT_EINVAL, "null argument");
    for (std::size_t k = 0; k < T; ++k)
    {
        if (counts[k] > 0 && (!x[k] || !y[k] || !z[k]))
            throw Error(AT_EINVAL, "adaptive_octree: coordinate device pointers must not be null");
        if (counts[k] > std::size_t(std::numeric_limits<std::uint32_t>::max()))
            throw Error(AT_EINVAL, "adaptive_octree: particle count exceeds 32-bit permutation index range");
    }
    const unsigned md = c.max_depth >= 0 ? unsigned(c.max_depth) : kOctLevels;
    return new OctreeBatch<Real>(T, reinterpret_cast<const Real *const *>(x), reinterpret_cast<const Real *const *>(y),
                                 reinterpret_cast<const Real *const *>(z), counts, c, md, it);
}

// ---------------------------------------------------------------------------
// Groups: members that keep their own buffers
// ---------------------------------------------------------------------------
// A member's segment of the group's sorted keys and the copy blocks that
// write it to the member's P and Perm.
struct SegOut
{
    Key *P;
    std::uint32_t *perm;
    std::uint64_t begin;      // first key of the segment in the group arrays
    std::uint32_t n;
    std::uint32_t firstBlock; // first copy block of the member
};
constexpr int kCopyThreads = 256, kCopyItems = 8, kCopyChunk = kCopyThreads * kCopyItems;

__global__ __launch_bounds__(kCopyThreads) void scatterSegmentsKernel(const SegOut *segs, int count, const Key *P,
                                                                       const std::uint32_t *perm)
{
    __shared__ int member;
    if (threadIdx.x == 0)
    {
        // the last member whose first block is at most this block
        int lo = 0, hi = count - 1;
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if (segs[mid].firstBlock <= blockIdx.x) lo = mid;
            else hi = mid - 1;
        }
        member = lo;
    }
    __syncthreads();
    const SegOut s = segs[member];
    const std::size_t base = std::size_t(blockIdx.x - s.firstBlock) * kCopyChunk;
#pragma unroll
    for (int j = 0; j < kCopyItems; ++j)
    {
        const std::size_t i = base + std::size_t(j) * kCopyThreads + threadIdx.x;
        if (i < s.n)
        {
            s.P[i] = P[s.begin + i];
            s.perm[i] = perm[s.begin + i];
        }
    }
}

__global__ void gatherStatesKernel(const TreeDev *trees, int count, DevState *out)
{
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k < count) out[k] = *trees[k].st;
}

__global__ void scatterStatesKernel(const DevState *in, const TreeDev *trees, int count)
{
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k < count) *trees[k].st = in[k];
}

// A table written in pinned host memory and copied to the device in stream
// order. Growing it waits for the stream (kernels may still read it).
template <class T>
struct Staged
{
    T *d = nullptr, *h = nullptr;
    std::size_t cap = 0;
    hipEvent_t copied = nullptr; // after the last copy out of h
    ~Staged() { release(); }
    void release()
    {
        if (d) (void)hipFree(d);
        if (h) (void)hipHostFree(h);
        if (copied) (void)hipEventDestroy(copied);
        d = h = nullptr;
        copied = nullptr;
        cap = 0;
    }
    T *host(std::size_t n, hipStream_t s)
    {
        if (n > cap)
        {
            syncStream(s);
            release();
            cap = n;
            AT_HIP(hipMalloc(&d, n * sizeof(T)));
            AT_HIP(hipHostMalloc(&h, n * sizeof(T)));
            AT_HIP(hipEventCreateWithFlags(&copied, hipEventDisableTiming));
        }
        else if (copied)
            AT_HIP(hipEventSynchronize(copied));
        return h;
    }
    void upload(hipStream_t s, std::size_t n)
    {
        if (!n) return;
        AT_HIP(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
        AT_HIP(hipEventRecord(copied, s));
    }
};

// Members of one class larger than this, or groups beyond this many
// particles, run on their own: launch overhead is what a group saves, and
// the workspace grows with the group.
constexpr std::size_t kGroupMaxTree = std::size_t(1) << 22;
constexpr std::size_t kGroupMaxTotal = std::size_t(1) << 24;
} // namespace

struct GroupRunner::Impl
{
    hipStream_t s = nullptr;
    SegmentedSort sort;
    DevBuf<Key> P;
    DevBuf<std::uint32_t> perm;
    Staged<SegCoords<float>> coords4;
    Staged<SegCoords<double>> coords8;
    Staged<SegOut> outs;
    SetTables tab;
    DevBuf<std::uint32_t> planStart, passLog;
    std::vector<std::uint32_t> logHost;
    // gathered member states: device [round], pinned host copy
    DevBuf<DevState> states;
    DevState *statesH = nullptr;
    std::size_t statesHCap = 0;
    struct Pending
    {
        int realBytes;
        std::vector<TreeBase *> members;
        std::size_t offset; // of the members' states in `states`
    };
    std::vector<Pending> pending;
    std::size_t pendingStates = 0;

    ~Impl()
    {
        (void)syncStreamStatus(s);
        if (statesH) (void)hipHostFree(statesH);
    }

    DevState *hostStates(std::size_t n)
    {
        if (n > statesHCap)
        {
            if (statesH) AT_HIP(hipHostFree(statesH));
            statesH = nullptr;
            AT_HIP(hipHostMalloc(&statesH, n * sizeof(DevState)));
            statesHCap = n;
        }
        return statesH;
    }
    template <class Real>
    Staged<SegCoords<Real>> &coordTable()
    {
        if constexpr (sizeof(Real) == 4) return coords4;
        else return coords8;
    }

    template <class Real>
    std::vector<impl::OctreeImpl<Real> *> cast(const std::vector<TreeBase *> &members) const
    {
        using M = impl::OctreeImpl<Real>;
        std::vector<M *> out;
        out.reserve(members.size());
        for (TreeBase *t : members)
        {
            M *m = dynamic_cast<M *>(t);
            if (!m) throw Error(AT_EINVAL, "group members are octrees of one precision");
            if (m->stream() != s) throw Error(AT_EINVAL, "group members run on the group's stream");
            if (m->particles() == 0) throw Error(AT_EINVAL, "group members hold particles");
            if (!out.empty())
            {
                const M &a = *out[0];
                const OctParams p = a.parameters(), q = m->parameters();
                const GridBox<Real> &g = a.gridBox(), &h = m->gridBox();
                const bool same = p.bucket == q.bucket && p.nf == q.nf && p.hilbert == q.hilbert &&
                                  a.depthLimit() == m->depthLimit() && a.updateIterations() == m->updateIterations() &&
                                  g.xmin == h.xmin && g.ymin == h.ymin && g.zmin == h.zmin && g.mx == h.mx &&
                                  g.my == h.my && g.mz == h.mz && g.cx == h.cx && g.cy == h.cy && g.cz == h.cz;
                if (!same) throw Error(AT_EINVAL, "group members share one configuration");
            }
            out.push_back(m);
        }
        return out;
    }

    template <class M>
    std::uint32_t blocksFor(Shape sh, M &m, std::size_t leaves)
    {
        switch (sh)
        {
        case kClear: return clearBlocksFor(m.device().statusWords);
        case kTiles: return tileBlocksFor(leaves);
        case kLeaves: return leafBlocksFor(leaves);
        case kIndex: return indexBlocksFor(m.particles());
        case kTileFirst: return tileFirstBlocksFor(m.particles());
        case kTilesOfP: return tileBlocksOfP(m.particles());
        default: return m.parameters().nf ? nfBlocksFor(leaves) : 0u;
        }
    }

    // The members' keys sorted together, each segment written to its
    // member's P and Perm.
    template <class Real>
    void sortInto(const std::vector<impl::OctreeImpl<Real> *> &m, bool hostPoll)
    {
        const std::size_t G = m.size();
        std::vector<std::size_t> counts(G);
        for (std::size_t k = 0; k < G; ++k) counts[k] = m[k]->particles();
        sort.plan(counts, s);
        const std::size_t total = sort.total();
        P.ensure(total, 0, s);
        perm.ensure(total, 0, s);
        Staged<SegCoords<Real>> &ct = coordTable<Real>();
        SegCoords<Real> *hc = ct.host(G, s);
        for (std::size_t k = 0; k < G; ++k) hc[k] = m[k]->coordinates();
        ct.upload(s, G);
        KeygenParams<Real> kp{};
        kp.grid = m[0]->gridBox();
        const KeyKind kind = m[0]->parameters().hilbert ? KeyKind::Hilbert : KeyKind::Morton;
        sort.keygenAndSort<Real>(s, kind, kp, ct.d, P.p, perm.p, hostPoll);
        SegOut *ho = outs.host(G, s);
        std::uint32_t blocks = 0;
        for (std::size_t k = 0; k < G; ++k)
        {
            ho[k] = SegOut{m[k]->sortedKeys(), m[k]->sortedPerm(), std::uint64_t(sort.segBegin(k)),
                           std::uint32_t(counts[k]), blocks};
            blocks += std::uint32_t((counts[k] + kCopyChunk - 1) / kCopyChunk);
        }
        outs.upload(s, G);
        hipLaunchKernelGGL(scatterSegmentsKernel, dim3(blocks), dim3(kCopyThreads), 0, s, outs.d, int(G), P.p,
                           perm.p);
        AT_HIP(hipGetLastError());
    }

    // TreeDev table of the members (adjusted by `adjust`) and block maps
    // from `leaves(k)`.
    template <class M, class A, class L>
    void stage(const std::vector<M *> &m, A adjust, L leaves, TreeSet set[kShapes])
    {
        if (m.size() > tab.cap)
        {
            syncStream(s); // earlier launches may still read the tables
            tab.reserve(m.size());
        }
        tab.waitWritable();
        for (std::size_t k = 0; k < m.size(); ++k)
        {
            TreeDev t = m[k]->device();
            adjust(k, t);
            tab.hTrees[k] = t;
        }
        tab.upload(s, m.size(), [&](Shape sh, std::size_t k) { return blocksFor(sh, *m[k], leaves(k)); }, set);
    }

    void gather(std::size_t G, std::size_t offset)
    {
        states.ensure(offset + G, offset, s);
        hipLaunchKernelGGL(gatherStatesKernel, dim3(unsigned((G + 63) / 64)), dim3(64), 0, s, tab.dTrees, int(G),
                           states.p + offset);
        AT_HIP(hipGetLastError());
    }

    // OctreeBatch::build over members with their own buffers.
    template <class Real>
    void build(const std::vector<TreeBase *> &members)
    {
        if (!pending.empty()) throw Error(AT_ELOGIC, "group build with updates pending");
        const auto m = cast<Real>(members);
        const std::size_t G = m.size();
        if (!G) return;
        syncStream(s);
        sortInto<Real>(m, false);
        for (auto *t : m)
        {
            t->buildBegin();
            t->reserveDeviceBuild(t->deviceBuildCapacity());
        }
        const unsigned maxDepth = m[0]->depthLimit();
        const OctParams p = m[0]->parameters();
        const std::size_t logStride = maxDepth + 1;
        passLog.ensure(G * logStride + 1, 0, s);
        TreeSet set[kShapes];
        stage(m, [&](std::size_t k, TreeDev &t) {
            t.capacity = m[k]->deviceBuildLimit();
            t.passLog = passLog.p + k * logStride;
        }, [](std::size_t) { return std::size_t(1); }, set);
        multiBuildSearchIndex(s, set[kIndex]);
        multiOctInit(s, set[kClear]);
        planStart.ensure(std::size_t(kPlanShapes) * (G + 1), 0, s);
        const TreeDev *dt = tab.dTrees;
        const int count = int(G);
        const TreeSet tiles = plannedSet(dt, count, planStart.p, kPlanTiles);
        const TreeSet proc = plannedSet(dt, count, planStart.p, kPlanProcess);
        const TreeSet commit = plannedSet(dt, count, planStart.p, kPlanCommit);
        const TreeSet fresh = plannedSet(dt, count, planStart.p, kPlanLeaves);
        const TreeSet nfWork = plannedSet(dt, count, planStart.p, kPlanNf);
        for (unsigned pass = 0; pass < maxDepth; ++pass)
        {
            multiPlan(s, dt, count, 0, p.nf, planStart.p);
            multiOctDecideScan(s, tiles, p);
            multiPlan(s, dt, count, 1, p.nf, planStart.p);
            multiOctProcess(s, proc, p);
            multiCommit(s, commit, false);
            multiOctRefresh(s, fresh, nfWork, p, false);
        }
        // the one read-back: states and the committed passes' leaf counts
        gather(G, 0);
        DevState *sh = hostStates(G);
        logHost.resize(G * logStride);
        AT_HIP(hipMemcpyAsync(sh, states.p, G * sizeof(DevState), hipMemcpyDeviceToHost, s));
        AT_HIP(hipMemcpyAsync(logHost.data(), passLog.p, G * logStride * sizeof(std::uint32_t),
                              hipMemcpyDeviceToHost, s));
        syncStream(s);
        std::vector<std::size_t> stalled;
        std::vector<unsigned> decided(G, 0);
        for (std::size_t k = 0; k < G; ++k)
        {
            const unsigned passes = std::min(sh[k].passes, maxDepth);
            for (unsigned j = 0; j < passes; ++j) m[k]->replayBuildPass(logHost[k * logStride + j]);
            decided[k] = passes;
            if (sh[k].needsResize) stalled.push_back(k);
        }
        std::vector<DevState> stalledStates;
        for (std::size_t k : stalled) stalledStates.push_back(sh[k]);
        for (std::size_t k = 0; k < G; ++k)
        {
            if (sh[k].needsResize) continue;
            m[k]->buildEnd(false);
            impl::OctreeImpl<Real>::finishState(sh[k], m[k]->buildLeaves());
        }
        AT_HIP(hipMemcpyAsync(states.p, sh, G * sizeof(DevState), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(scatterStatesKernel, dim3(unsigned((G + 63) / 64)), dim3(64), 0, s, states.p, tab.dTrees,
                           count);
        AT_HIP(hipGetLastError());
        for (std::size_t k = 0; k < G; ++k)
        {
            if (sh[k].needsResize) continue;
            if (m[k]->configuredViews().any()) m[k]->buildViews(m[k]->configuredViews());
            (void)m[k]->adoptState(sh[k]);
        }
        // completes the buffers' growth copies and the uploads above
        syncStream(s);
        // members whose pass outgrew the reservation finish on their own
        for (std::size_t i = 0; i < stalled.size(); ++i) m[stalled[i]]->continueBuild(stalledStates[i], decided[stalled[i]]);
    }

    // OctreeBatch::update over members with their own buffers; the states
    // are gathered for readBack.
    template <class Real>
    void launchUpdate(const std::vector<TreeBase *> &members)
    {
        const auto m = cast<Real>(members);
        const std::size_t G = m.size();
        if (!G) return;
        for (auto *t : m)
            if (!t->isBuilt()) throw Error(AT_ELOGIC, "adaptive_octree: call build() before update()");
        sortInto<Real>(m, true);
        TreeSet set[kShapes];
        stage(m, [](std::size_t, TreeDev &) {}, [&](std::size_t k) { return std::size_t(m[k]->device().capacity); },
              set);
        const OctParams p = m[0]->parameters();
        multiBuildSearchIndex(s, set[kIndex]);
        multiReset(s, set[kClear]);
        // every member recounts all leaves: by tiles of P for small LeafCount leaves
        if (!p.nf && p.bucket <= kTileCountBucket)
            multiLeafCountsByTiles(s, set[kTileFirst], set[kTilesOfP], set[kLeaves]);
        else
            multiOctRefresh(s, set[kLeaves], set[kNf], p, true);
        for (unsigned i = 0; i < m[0]->updateIterations(); ++i)
        {
            multiOctDecideScan(s, set[kTiles], p);
            multiOctProcess(s, set[kLeaves], p);
            multiCommit(s, set[kClear], false);
            multiOctRefresh(s, set[kLeaves], set[kNf], p, false);
        }
        for (auto *t : m)
            if (t->configuredViews().any()) t->buildViews(t->configuredViews());
        gather(G, pendingStates);
        pending.push_back(Pending{int(sizeof(Real)), members, pendingStates});
        pendingStates += G;
    }

    template <class Real>
    void finish(const Pending &pd, std::vector<char> &flags)
    {
        const auto m = cast<Real>(pd.members);
        flags.assign(m.size(), 0);
        for (std::size_t k = 0; k < m.size(); ++k) flags[k] = m[k]->adoptState(statesH[pd.offset + k]) ? 1 : 0;
    }
};

GroupRunner::GroupRunner(hipStream_t s) : impl_(new Impl)
{
    impl_->s = s;
}

GroupRunner::~GroupRunner() = default;

bool GroupRunner::groupable(int type, const at_config &c, std::size_t n)
{
    return type == AT_TREE_OCTREE && n > 0 && n <= kGroupMaxTree && c.max_particles_per_leaf > 0;
}

std::size_t GroupRunner::maxParticles() { return kGroupMaxTotal; }

void GroupRunner::build(const std::vector<TreeBase *> &members)
{
    if (members.empty()) return;
    if (dynamic_cast<impl::OctreeImpl<float> *>(members[0])) impl_->build<float>(members);
    else impl_->build<double>(members);
}

void GroupRunner::launchUpdate(const std::vector<TreeBase *> &members)
{
    if (members.empty()) return;
    if (dynamic_cast<impl::OctreeImpl<float> *>(members[0])) impl_->launchUpdate<float>(members);
    else impl_->launchUpdate<double>(members);
}

void GroupRunner::readBack()
{
    Impl &g = *impl_;
    if (!g.pendingStates) return;
    DevState *h = g.hostStates(g.pendingStates);
    AT_HIP(hipMemcpyAsync(h, g.states.p, g.pendingStates * sizeof(DevState), hipMemcpyDeviceToHost, g.s));
}

void GroupRunner::finishUpdates(std::vector<std::vector<char>> &needsResize)
{
    Impl &g = *impl_;
    needsResize.assign(g.pending.size(), {});
    for (std::size_t i = 0; i < g.pending.size(); ++i)
    {
        if (g.pending[i].realBytes == 4) g.finish<float>(g.pending[i], needsResize[i]);
        else g.finish<double>(g.pending[i], needsResize[i]);
    }
    discardUpdates();
}

void GroupRunner::discardUpdates()
{
    impl_->pending.clear();
    impl_->pendingStates = 0;
}

BatchBase::~BatchBase() = default;

BatchBase *createBatch(int realBytes, int type, std::size_t T, const void *const *x, const void *const *y,
                       const void *const *z, const std::size_t *counts, const at_config &c)
{
    if (realBytes == 4) return makeBatch<float>(type, T, x, y, z, counts, c);
    if (realBytes == 8) return makeBatch<double>(type, T, x, y, z, counts, c);
    throw Error(AT_EINVAL, "real_bytes must be 4 or 8");
}
} // namespace at

// ---------------------------------------------------------------------------
// C ABI of batches
// ---------------------------------------------------------------------------
struct at_batch
{
    std::unique_ptr<at::BatchBase> impl;
};

namespace
{
template <class F>
int guard(F &&f)
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

at::BatchBase &batchOf(at_batch *b)
{
    if (!b || !b->impl) throw at::Error(AT_EINVAL, "null batch");
    return *b->impl;
}
} // namespace

extern "C" {
int at_batch_create(int realBytes, int type, size_t numTrees, const void *const *x, const void *const *y,
                    const void *const *z, const size_t *counts, const at_config *config, at_batch **out)
{
    return guard([&] {
        if (!config || !out) throw at::Error(AT_EINVAL, "null argument");
        auto b = std::make_unique<at_batch>();
        b->impl.reset(at::createBatch(realBytes, type, numTrees, x, y, z, counts, *config));
        *out = b.release();
    });
}
int at_batch_destroy(at_batch *b)
{
    return guard([&] { delete b; });
}
int at_batch_build(at_batch *b)
{
    return guard([&] { batchOf(b).build(); });
}
int at_batch_update(at_batch *b)
{
    return guard([&] { batchOf(b).update(); });
}
int at_batch_needs_resize(at_batch *b, int *flags, size_t *count)
{
    return guard([&] {
        const std::size_t n = batchOf(b).needsResize(flags);
        if (count) *count = n;
    });
}
int at_batch_resize_buffers(at_batch *b)
{
    return guard([&] { batchOf(b).resizeBuffers(); });
}
int at_batch_size(at_batch *b, size_t *n)
{
    return guard([&] {
        if (!n) throw at::Error(AT_EINVAL, "null output pointer");
        *n = batchOf(b).size();
    });
}
int at_batch_tree_buffer(at_batch *b, size_t tree, int id, const void **d, size_t *s, size_t *c)
{
    return guard([&] {
        if (!d || !s || !c) throw at::Error(AT_EINVAL, "null output pointer");
        batchOf(b).tree(tree).buffer(id, d, s, c);
    });
}
int at_batch_tree_scalars(at_batch *b, size_t tree, at_scalars *o)
{
    return guard([&] {
        if (!o) throw at::Error(AT_EINVAL, "null output pointer");
        batchOf(b).tree(tree).scalars(o);
    });
}
int at_batch_graph_capture_update(at_batch *b, void *stream, void **exec)
{
    return guard([&] {
        if (!exec) throw at::Error(AT_EINVAL, "null output pointer");
        at::BatchBase &batch = batchOf(b);
        hipStream_t s = batch.stream();
        if (!s) throw at::Error(AT_EINVAL, "graph capture needs a batch created on a non-default stream");
        if (stream && hipStream_t(stream) != s)
            throw at::Error(AT_EINVAL, "capture stream must be the batch's configured stream");
        hipGraph_t g = nullptr;
        AT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeGlobal));
        try
        {
            batch.update();
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
}
