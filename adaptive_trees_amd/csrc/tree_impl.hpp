// tree_impl.hpp -- host control of one octree / binary tree (the objects
// behind the C ABI and the members of a batch).
//
// Octree<Real> follows src/octree/octree.cu, KDTree3D<Real> follows
// src/kdtree3d/kdtree3d.cu: same call protocol (build, capturable update,
// doBuffersNeedResize, resizeBuffers, computeViews), same capacity rules and
// the same device-resident RebalanceState, driven by the kernels of
// tree_kernels.hip, sort.hip and views.hip on the caller's stream.
#pragma once

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>

#include "engine.hpp"

namespace at
{
namespace impl
{
constexpr unsigned kGrowthDivisor = 4; // +25% (config.hpp:62)

// Box<T> of the reference (sfc_keys.cuh:420-466), evaluated on the host in T.
template <class T>
GridBox<T> makeGridBox(const at_box &b)
{
    const T xmin = T(b.xmin), xmax = T(b.xmax), ymin = T(b.ymin), ymax = T(b.ymax), zmin = T(b.zmin),
            zmax = T(b.zmax);
    const T ilx = T(1) / (xmax - xmin), ily = T(1) / (ymax - ymin), ilz = T(1) / (zmax - zmin);
    const T lx = xmax - xmin, ly = ymax - ymin, lz = zmax - zmin;
    const T mxy = lx > ly ? lx : ly;
    const T maxDim = mxy > lz ? mxy : lz;
    const T bias = T(0.5849625007211563);
    const unsigned bx = kOctLevels - static_cast<unsigned>(std::floor(std::log2(maxDim / lx) + bias));
    const unsigned by = kOctLevels - static_cast<unsigned>(std::floor(std::log2(maxDim / ly) + bias));
    const unsigned bz = kOctLevels - static_cast<unsigned>(std::floor(std::log2(maxDim / lz) + bias));
    GridBox<T> g;
    g.xmin = xmin; g.ymin = ymin; g.zmin = zmin;
    g.cx = int(1u << bx); g.cy = int(1u << by); g.cz = int(1u << bz);
    g.mx = T(unsigned(g.cx)) * ilx;
    g.my = T(unsigned(g.cy)) * ily;
    g.mz = T(unsigned(g.cz)) * ilz;
    return g;
}

// makeBinaryGeometry (binary_node.cuh:17-40) plus the fast-path split of the
// axis schedule into a prefix and a period-3 cycle.
template <class T>
BinaryGeom makeBinaryGeom(const at_box &box, unsigned maxDepth)
{
    if (maxDepth > kBinLevels) throw Error(AT_EINVAL, "Binary maxDepth exceeds key precision");
    long double len[3] = {static_cast<long double>(box.xmax) - box.xmin, static_cast<long double>(box.ymax) - box.ymin,
                          static_cast<long double>(box.zmax) - box.zmin};
    for (auto side : len)
        if (!(side > 0) || !std::isfinite(side)) throw Error(AT_EINVAL, "Tree box must have finite positive sides");
    BinaryGeom g{};
    for (unsigned d = 0; d < kBinLevels; ++d)
    {
        unsigned a = 0;
        if (len[1] > len[a]) a = 1;
        if (len[2] > len[a]) a = 2;
        g.axis[d] = std::uint8_t(a);
        len[a] *= 0.5L;
    }
    // the key is computed from the box stored in Real, widened to double
    const T lo[3] = {T(box.xmin), T(box.ymin), T(box.zmin)};
    const T hi[3] = {T(box.xmax), T(box.ymax), T(box.zmax)};
    for (int a = 0; a < 3; ++a)
    {
        g.lo[a] = double(lo[a]);
        g.len[a] = double(hi[a]) - double(lo[a]);
    }
    g.periodic = 1;
    g.prefix = int(kBinLevels);
    for (int p = 0; p + 3 <= int(kBinLevels); ++p)
    {
        const int a0 = g.axis[p], a1 = g.axis[p + 1], a2 = g.axis[p + 2];
        if (a0 == a1 || a1 == a2 || a0 == a2) continue;
        bool ok = true;
        for (int d = p; d < int(kBinLevels) && ok; ++d) ok = g.axis[d] == g.axis[p + (d - p) % 3];
        if (ok)
        {
            g.prefix = p;
            g.order[0] = a0; g.order[1] = a1; g.order[2] = a2;
            break;
        }
    }
    g.triples = (int(kBinLevels) - g.prefix + 2) / 3;
    for (int a = 0; a < 3; ++a) g.consumed[a] = 0;
    for (int d = 0; d < g.prefix; ++d) ++g.consumed[g.axis[d]];
    g.prefixAxis = g.prefix > 0 ? g.axis[0] : 0;
    for (int d = 0; d < g.prefix; ++d)
        if (g.axis[d] != g.axis[0]) g.prefixAxis = -1;
    g.fastDiv = 1;
    for (int a = 0; a < 3; ++a)
    {
        g.inv[a] = 1.0 / g.len[a]; // IEEE division: correctly rounded
        const double lo = std::ldexp(1.0, -1000), hi = std::ldexp(1.0, 1000);
        if (!(g.len[a] >= lo && g.len[a] <= hi && g.inv[a] >= lo && g.inv[a] <= hi)) g.fastDiv = 0;
    }
    // the branch-free form (binaryKeyFast)
    g.fast = g.fastDiv && g.prefix < int(kBinLevels) && (g.prefix == 0 || g.prefixAxis >= 0) ? 1 : 0;
    for (int i = 0; i < 3 && g.fast; ++i)
    {
        const int a = g.order[i];
        const int bits = g.triples + (g.prefix > 0 && a == g.prefixAxis ? g.prefix : 0);
        if (bits > 32) g.fast = 0;
        g.slotLo[i] = g.lo[a];
        g.slotLen[i] = g.len[a];
        g.slotInv[i] = g.inv[a];
        g.slotScale[i] = std::ldexp(1.0, bits);
        g.slotMax[i] = bits == 32 ? 0xffffffffu : (std::uint32_t(1) << bits) - 1;
    }
    return g;
}

inline std::size_t growBy25(std::size_t c) { return c + (c + kGrowthDivisor - 1) / kGrowthDivisor; }

// ---------------------------------------------------------------------------
// Octree<Real> (src/octree/octree.cu)
// ---------------------------------------------------------------------------
template <class Real>
class OctreeImpl final : public TreeBase
{
public:
    OctreeImpl(const Real *x, const Real *y, const Real *z, std::size_t n, const at_config &c, unsigned maxDepth,
               unsigned iterations, const MemberSlot *slot = nullptr)
        : TreeBase(hipStream_t(c.stream), slot ? slot->st : nullptr, slot ? slot->sth : nullptr), view_(3)
    {
        x_ = x; y_ = y; z_ = z; n_ = n;
        maxDepth_ = maxDepth;
        bucket_ = c.max_particles_per_leaf;
        iterations_ = iterations;
        nf_ = c.split_criterion == AT_NF_COUNT;
        hilbert_ = c.sfc_kind == AT_HILBERT;
        box_ = c.box;
        views_ = {c.views[0] != 0, c.views[1] != 0, c.views[2] != 0};
        prepared_ = views_;
        grid_ = makeGridBox<Real>(box_);
        if (slot)
        {
            Pp_ = slot->P;
            permp_ = slot->perm;
        }
        else
        {
            P_.ensure(n ? n : 1, 0, stream_);
            perm_.ensure(n ? n : 1, 0, stream_);
            sort_.allocate(n ? n : 1, true, stream_);
            Pp_ = P_.p;
            permp_ = perm_.p;
        }
        idx_.ensure(indexSize(n) + 1, 0, stream_);
        if (!nf_) tileFirst_.ensure(n / kBoundTile + 2, 0, stream_);
        vK_.resize(2);
        ensurePhysical(8, 0);
        syncStream(stream_);
    }

    // members release their memory as idle blocks: finish this tree's work first
    ~OctreeImpl() override { (void)syncStreamStatus(stream_); }

    void build() override
    {
        syncStream(stream_);
        computeParticleKeys();
        buildBegin();
        octInitRoot(stream_, dev());
        clearedStatus_ = status_.p;
        buildPasses(0);
        finishBuild();
    }

    // A group member whose device-driven build stopped at a pass beyond its
    // reservation: `s` is its device state after `decided` committed passes
    // (mirrored on the host); the remaining passes run as in build().
    void continueBuild(DevState s, unsigned decided)
    {
        s.needsResize = 0;
        s.converged = 0;
        *sth_ = s;
        AT_HIP(hipMemcpyAsync(st_, sth_, sizeof(DevState), hipMemcpyHostToDevice, stream_));
        clearedStatus_ = nullptr; // the rejected decision left its scan status behind
        buildPasses(decided);
        finishBuild();
    }

    // ---- members of a group keep their own buffers: the group sorts the
    // members' keys together and writes each member's segment to its P and
    // Perm, then drives every member through one launch per step ----
    SegCoords<Real> coordinates() const { return SegCoords<Real>{x_, y_, z_}; }
    std::size_t particles() const { return n_; }
    Key *sortedKeys() { return Pp_; }
    std::uint32_t *sortedPerm() { return permp_; }
    unsigned depthLimit() const { return maxDepth_; }
    unsigned updateIterations() const { return iterations_; }
    const GridBox<Real> &gridBox() const { return grid_; }
    // A state the group read back: the pinned host copy and the host sizes;
    // returns needsResize.
    bool adoptState(const DevState &s)
    {
        *sth_ = s;
        return applyState(s);
    }

    // ---- the build loop of octree.cu:159-180 in steps, so that a batch can
    // drive many trees through each step with one launch ----
    void buildBegin()
    {
        vK_.resize(2);
        activeBuffer_ = 0;
        buildL_ = 1;
        built_ = false;
    }
    std::size_t buildLeaves() const { return buildL_; }
    std::size_t buildNewLeaves() const { return buildNewL_; }
    // Physical buffers of the build passes follow the capacities of the
    // mirrored vectors (which grow with 25% headroom), not the exact leaf
    // counts: a pass that adds a few leaves then reuses the buffers instead
    // of reallocating every one of them (a new pool block each time).
    void buildPrePass() { ensureOps(std::max(buildL_ + 1, vK_.cap)); }
    // After the decision pass: false when the tree did not change (done).
    bool buildAfterDecide(const DevState &s)
    {
        if (!s.changed) return false;
        buildNewL_ = std::size_t(s.newNumLeaves);
        // the reference grows the target buffers with 25% headroom
        vTmp_.resizeHeadroom(buildNewL_ + 1);
        ensureInactive(std::max(buildNewL_ + 1, vTmp_.cap));
        return true;
    }
    void buildAfterProcess()
    {
        std::swap(vK_, vTmp_);
        activeBuffer_ ^= 1;
        buildL_ = buildNewL_;
    }
    // ---- a build whose passes run on the device without host read-backs
    // (batches): the buffers are reserved up front for `deviceBuildCapacity`
    // leaves, a pass that would exceed them is rejected on the device (the
    // tree stays as it was), and the host replays the committed passes'
    // leaf counts afterwards.
    // LeafCount: level jumps make at most 4096 leaves per 512 * bucket
    // particles, plain splits fewer; near field: up to one leaf per two
    // particles. A tree that needs more continues with host-driven passes.
    std::size_t deviceBuildCapacity() const
    {
        const std::size_t base = nf_ ? n_ / 2 : 8 * (n_ / std::max(bucket_, 1u));
        return std::max<std::size_t>(base + 4096 + 64, vK_.cap);
    }
    void reserveDeviceBuild(std::size_t leaves) { ensurePhysical(leaves, 0); }
    // The capacity a device pass may fill: the reserved halves and scan buffers.
    long deviceBuildLimit() const
    {
        std::size_t c = std::min(K_[0].cap, K_[1].cap);
        c = c ? c - 1 : 0;
        c = std::min({c, N_[0].cap, N_[1].cap, offs_.cap ? offs_.cap - 1 : 0});
        if (nf_) c = std::min({c, NF_[0].cap, NF_[1].cap, gcm_[0].cap, gcm_[1].cap, nfDirty_.cap, groupDirty_.cap,
                               tasks_.cap / 2, groups_.cap ? (groups_.cap - 1) * 8 : 0});
        else c = std::min(c, countDirty_.cap);
        return long(c);
    }
    // buildAfterDecide + buildAfterProcess of one committed device pass, on
    // the mirrored capacities only.
    // Large LeafCount trees with small leaves recount by tiles of P (nearly
    // every line of P would be searched anyway: 16M Plummer particles, bucket
    // 64: 0.84 vs 0.90 ms per update); below 4M particles the three launches
    // cost more than the searches. Near-field trees refresh N with NF.
    bool fullCountsByTiles() const { return !nf_ && bucket_ <= kTileCountBucket && n_ >= (std::size_t(1) << 22); }
    void replayBuildPass(std::size_t newLeaves)
    {
        buildNewL_ = newLeaves;
        vTmp_.resizeHeadroom(newLeaves + 1);
        std::swap(vK_, vTmp_);
        activeBuffer_ ^= 1;
        buildL_ = newLeaves;
    }
    // Half 0 becomes the active one, as the reference keeps the tree in K_d;
    // capacity from the vector capacities reached (octree.cu:182-241,453-457).
    // sync = false leaves the stream synchronisation to the caller.
    void buildEnd(bool sync)
    {
        if (activeBuffer_) swapHalves();
        activeBuffer_ = 0;
        const std::size_t L = buildL_;
        const std::size_t cap =
            std::max({vK_.cap - 1, vTmp_.size == 0 ? std::size_t(0) : vTmp_.cap - 1, L + (L + 3) / 4});
        prepareGraphBuffers(cap, L, sync);
        built_ = true;
    }
    // finishBuildStateKernel (octree.cu:40-47) applied to a state on the host.
    static void finishState(DevState &s, std::size_t L)
    {
        s.numLeaves = s.newNumLeaves = int(L);
        s.changed = s.converged = s.needsResize = 0;
        s.numNodes = 0;
        s.activeBuffer = 0;
    }
    const ViewFlags &configuredViews() const { return views_; }
    void buildViews(const ViewFlags &v) { buildTreeView(v); }
    // Host sizes from a state read back by the caller; returns needsResize.
    bool applyState(const DevState &s)
    {
        activeBuffer_ = s.activeBuffer;
        hostLeaves_ = std::size_t(s.numLeaves);
        vK_.resize(hostLeaves_ + 1);
        vTmp_.resize(hostLeaves_ + 1);
        numActive_ = s.numActiveParticles;
        view_.syncSizes(hostLeaves_, numActive_);
        numNodes_ = std::size_t(s.numNodes);
        return s.needsResize != 0;
    }
    // resizeBuffers after applyState, without clearing the device flag or
    // synchronising (the caller does both).
    void growForResize() { prepareGraphBuffers(growBy25(leafCapacity_), hostLeaves_, false); }
    bool isBuilt() const { return built_; }
    TreeDev device() { return dev(); }
    OctParams parameters() const { return params(); }

    void update() override
    {
        if (!built_) throw Error(AT_ELOGIC, "adaptive_octree: call build() before update()");
        computeParticleKeys();
        const TreeDev t = dev();
        const OctParams p = params();
        updateReset(stream_, t);
        // the particles moved: count every leaf (small leaves: one read of P)
        if (fullCountsByTiles()) leafCountsByTiles(stream_, t);
        else octRefresh(stream_, t, p, true, leafCapacity_);
        for (unsigned i = 0; i < iterations_; ++i) octPass(stream_, t, p, long(leafCapacity_), leafCapacity_);
        buildTreeView(views_);
        AT_HIP(hipGetLastError());
    }

    void computeViews(const ViewFlags *v) override
    {
        if (!built_) throw Error(AT_ELOGIC, "adaptive_octree: call build() before computeViews()");
        buildTreeView(v ? *v : views_);
    }

    bool needsResize() override
    {
        if (!built_) return false;
        readState();
        return applyState(*sth_);
    }

    void resizeBuffers() override
    {
        if (!built_) throw Error(AT_ELOGIC, "adaptive_octree: call build() before resizeBuffers()");
        (void)needsResize();
        prepareGraphBuffers(growBy25(leafCapacity_), hostLeaves_);
        AT_HIP(hipMemsetAsync(&st_->needsResize, 0, sizeof(int), stream_));
        syncStream(stream_);
    }

    void buffer(int id, const void **d, std::size_t *s, std::size_t *c) override
    {
        const int ab = activeBuffer_;
        switch (id)
        {
        case AT_BUF_PARTICLE_KEYS: *d = Pp_; *s = *c = n_; return;
        case AT_BUF_PERM: *d = permp_; *s = *c = n_; return;
        case AT_BUF_CORNERSTONE: *d = K_[ab].p; *s = hostLeaves_ + 1; *c = leafCapacity_ + 1; return;
        case AT_BUF_LEAF_COUNTS: *d = N_[ab].p; *s = hostLeaves_; *c = leafCapacity_; return;
        case AT_BUF_NEAR_FIELD_COUNTS: *d = NF_[ab].p; *s = hostLeaves_; *c = leafCapacity_; return;
        default: viewBuffer(view_, id, d, s, c);
        }
    }

    void scalars(at_scalars *o) override
    {
        o->num_particles = n_;
        o->num_active_particles = numActive_;
        o->num_nodes = numNodes_;
        o->num_leaves = hostLeaves_;
        o->leaf_capacity = leafCapacity_;
        o->tree_type = AT_TREE_OCTREE;
        o->split_criterion = nf_ ? AT_NF_COUNT : AT_LEAF_COUNT;
        o->sfc_kind = hilbert_ ? AT_HILBERT : AT_MORTON;
        o->real_bytes = int(sizeof(Real));
        o->active_buffer = activeBuffer_;
        o->update_iterations = int(iterations_);
        o->max_depth = int(maxDepth_);
        o->max_particles_per_leaf = bucket_;
    }

    unsigned maxAchievedDepth() override { return view_.maxAchievedDepth(stream_); }

    static void viewBuffer(const ViewEngine &v, int id, const void **d, std::size_t *s, std::size_t *c)
    {
        ViewEngine::Out o;
        switch (id)
        {
        case AT_BUF_SFC_BOX_INDEX: o = ViewEngine::BoxIndex; break;
        case AT_BUF_HAS_BOX_SPLIT: o = ViewEngine::HasSplit; break;
        case AT_BUF_BOX_DEPTH: o = ViewEngine::Depth; break;
        case AT_BUF_PARENT_INDEX: o = ViewEngine::Parent; break;
        case AT_BUF_CHILD_INDEX: o = ViewEngine::Child; break;
        case AT_BUF_LEVEL_OFFSET: o = ViewEngine::LevelOffset; break;
        case AT_BUF_PARTICLE_BEGIN_INDEX: o = ViewEngine::Begin; break;
        case AT_BUF_PARTICLE_COUNTS: o = ViewEngine::Counts; break;
        case AT_BUF_SFC_PARTICLE_INDEX: o = ViewEngine::Mapping; break;
        default: throw Error(AT_EINVAL, "unknown buffer id");
        }
        v.buffer(o, d, s, c);
    }

private:
    OctParams params() const { return OctParams{bucket_, nf_, hilbert_}; }

    // Host-driven passes from decision `first` on (octree.cu:159-180): each
    // decision is read back; the loop ends at the pass that changes nothing.
    void buildPasses(unsigned first)
    {
        const OctParams p = params();
        for (unsigned i = first; i < maxDepth_; ++i)
        {
            buildPrePass();
            // the scan status and ticket are zero after the root's init and
            // after every commit; a status buffer regrown since is cleared
            if (status_.p != clearedStatus_)
            {
                AT_HIP(hipMemsetAsync(status_.p, 0, status_.cap * sizeof(std::uint64_t), stream_));
                AT_HIP(hipMemsetAsync(ticket_.p, 0, sizeof(std::uint32_t), stream_));
                clearedStatus_ = status_.p;
            }
            TreeDev t = dev();
            octDecideScan(stream_, t, p, LONG_MAX, buildL_);
            readState();
            if (!buildAfterDecide(*sth_)) break;
            t = dev();
            octProcessCommit(stream_, t, p, buildL_);
            clearedStatus_ = status_.p;
            octRefresh(stream_, t, p, false, buildNewL_);
            buildAfterProcess();
        }
    }
    void finishBuild()
    {
        buildEnd(true);
        finishBuildState(buildL_);
        buildTreeView(views_);
        (void)needsResize();
    }

    TreeDev dev()
    {
        TreeDev t;
        for (int h = 0; h < 2; ++h)
        {
            t.K[h] = K_[h].p;
            t.N[h] = N_[h].p;
            t.NF[h] = nf_ ? NF_[h].p : nullptr;
            t.gcm[h] = nf_ ? gcm_[h].p : nullptr;
        }
        t.nfDirty = nf_ ? nfDirty_.p : nullptr;
        t.countDirty = nf_ ? nullptr : countDirty_.p;
        t.groupDirty = nf_ ? groupDirty_.p : nullptr;
        t.offs = offs_.p;
        t.scanStatus = status_.p;
        t.scanTicket = ticket_.p;
        t.nfTasks = nf_ ? tasks_.p : nullptr;
        t.nfGroups = nf_ ? groups_.p : nullptr;
        t.nfCount = nf_ ? taskCount_.p : nullptr;
        t.statusWords = status_.cap;
        t.st = st_;
        t.P = Pp_;
        t.Pidx = idx_.p;
        t.tileFirst = nf_ ? nullptr : tileFirst_.p;
        t.n = n_;
        t.capacity = long(leafCapacity_);
        return t;
    }

    void computeParticleKeys() // octree.cu:285-307
    {
        KeygenParams<Real> kp{};
        kp.x = x_; kp.y = y_; kp.z = z_;
        kp.grid = grid_;
        keygenAndSort<Real>(stream_, hilbert_ ? KeyKind::Hilbert : KeyKind::Morton, kp, n_, Pp_, permp_, sort_);
        buildSearchIndex(stream_, dev());
    }

    // physical buffers: both halves hold `leaves` entries (+1 for K)
    void ensurePhysical(std::size_t leaves, std::size_t keep)
    {
        for (int h = 0; h < 2; ++h)
        {
            K_[h].ensure(leaves + 1, keep ? keep + 1 : 0, stream_);
            N_[h].ensure(leaves, keep, stream_);
            if (nf_)
            {
                NF_[h].ensure(leaves, keep, stream_);
                gcm_[h].ensure(leaves, keep, stream_);
            }
        }
        if (!nf_) countDirty_.ensure(leaves, 0, stream_);
        if (nf_)
        {
            nfDirty_.ensure(leaves, 0, stream_);
            groupDirty_.ensure(leaves, 0, stream_);
            tasks_.ensure(2 * leaves, 0, stream_);
            groups_.ensure(leaves / 8 + 1, 0, stream_);
            taskCount_.ensure(2, 0, stream_);
        }
        ensureOps(leaves + 1);
        ticket_.ensure(1, 0, stream_);
    }
    void ensureOps(std::size_t items)
    {
        offs_.ensure(items, 0, stream_);
        status_.ensure(scanStatusWords(items), 0, stream_);
    }
    // build pass: the half that receives the new tree is the inactive one
    void ensureInactive(std::size_t newK)
    {
        const int h = activeBuffer_ ^ 1;
        K_[h].ensure(newK, 0, stream_);
        N_[h].ensure(newK, 0, stream_);
        if (!nf_) countDirty_.ensure(newK, 0, stream_);
        if (nf_)
        {
            NF_[h].ensure(newK, 0, stream_);
            gcm_[h].ensure(newK, 0, stream_);
            nfDirty_.ensure(newK, 0, stream_);
            groupDirty_.ensure(newK, 0, stream_);
            tasks_.ensure(2 * newK, 0, stream_);
            groups_.ensure(newK / 8 + 1, 0, stream_);
        }
    }
    void swapHalves()
    {
        K_[0].swap(K_[1]);
        N_[0].swap(N_[1]);
        NF_[0].swap(NF_[1]);
        gcm_[0].swap(gcm_[1]);
    }

    // prepareGraphBuffers (octree.cu:182-241)
    void prepareGraphBuffers(std::size_t capacity, std::size_t leaves, bool sync = true)
    {
        constexpr std::size_t maxLeaves = (std::size_t(std::numeric_limits<int>::max()) - 1) / 8 * 7 + 1;
        if (capacity > maxLeaves) throw Error(AT_ELENGTH, "adaptive_octree: buffer capacity exceeds node index range");
        leafCapacity_ = capacity;
        for (VecShape *v : {&vK_, &vTmp_})
        {
            if (v->cap < capacity + 1) v->reserve(capacity + 1);
            v->resize(capacity + 1);
        }
        // grow the physical halves; the active half keeps its live entries
        const int ab = activeBuffer_;
        K_[ab].ensure(capacity + 1, leaves + 1, stream_);
        N_[ab].ensure(capacity, leaves, stream_);
        if (nf_)
        {
            NF_[ab].ensure(capacity, leaves, stream_);
            gcm_[ab].ensure(capacity, leaves, stream_);
        }
        K_[ab ^ 1].ensure(capacity + 1, 0, stream_);
        N_[ab ^ 1].ensure(capacity, 0, stream_);
        if (!nf_) countDirty_.ensure(capacity, 0, stream_);
        if (nf_)
        {
            NF_[ab ^ 1].ensure(capacity, 0, stream_);
            gcm_[ab ^ 1].ensure(capacity, 0, stream_);
            nfDirty_.ensure(capacity, 0, stream_);
            groupDirty_.ensure(capacity, 0, stream_);
            tasks_.ensure(2 * capacity, 0, stream_);
            groups_.ensure(capacity / 8 + 1, 0, stream_);
        }
        ensureOps(capacity + 1);
        view_.prepare(capacity, n_, prepared_.structure, prepared_.counts, prepared_.mapping, stream_);
        vK_.resize(leaves + 1);
        vTmp_.resize(leaves + 1);
        view_.syncSizes(leaves, numActive_);
        hostLeaves_ = leaves;
        if (sync) syncStream(stream_);
    }

    void finishBuildState(std::size_t L) // finishBuildStateKernel (octree.cu:40-47)
    {
        AT_HIP(hipMemcpyAsync(sth_, st_, sizeof(DevState), hipMemcpyDeviceToHost, stream_));
        syncStream(stream_);
        DevState s = *sth_;
        s.numLeaves = s.newNumLeaves = int(L);
        s.changed = s.converged = s.needsResize = 0;
        s.numNodes = 0;
        s.activeBuffer = 0;
        *sth_ = s;
        AT_HIP(hipMemcpyAsync(st_, sth_, sizeof(DevState), hipMemcpyHostToDevice, stream_));
        syncStream(stream_);
    }

    void buildTreeView(const ViewFlags &v) // octree.cu:309-318
    {
        prepared_.structure |= v.structure;
        prepared_.counts |= v.counts;
        prepared_.mapping |= v.mapping;
        Key *const K[2] = {K_[0].p, K_[1].p};
        std::uint32_t *const N[2] = {N_[0].p, N_[1].p};
        view_.build(stream_, K, N, Pp_, n_, v.structure, v.counts, v.mapping, st_);
    }

    const Real *x_ = nullptr, *y_ = nullptr, *z_ = nullptr;
    std::size_t n_ = 0;
    unsigned maxDepth_ = 21, bucket_ = 64, iterations_ = 2;
    bool nf_ = false, hilbert_ = false;
    at_box box_{};
    GridBox<Real> grid_{};
    ViewFlags views_, prepared_;
    bool built_ = false;
    int activeBuffer_ = 0;
    std::size_t leafCapacity_ = 0, hostLeaves_ = 1, numActive_ = 0, numNodes_ = 0;
    std::size_t buildL_ = 1, buildNewL_ = 1;
    VecShape vK_, vTmp_;
    DevBuf<Key> P_;
    DevBuf<std::uint32_t> perm_;
    DevBuf<Key> idx_;               // sampled search index of P
    DevBuf<std::uint32_t> tileFirst_; // LeafCount: per tile of P, the first leaf boundary it owns (full counts)
    Key *Pp_ = nullptr;             // own P_ or the batch's segment
    std::uint32_t *permp_ = nullptr;
    DevBuf<Key> K_[2];
    DevBuf<std::uint32_t> N_[2], NF_[2];
    DevBuf<std::uint8_t> gcm_[2], nfDirty_, groupDirty_, countDirty_;
    DevBuf<std::int64_t> offs_;
    DevBuf<std::uint64_t> status_;
    std::uint64_t *clearedStatus_ = nullptr; // status_ as last cleared by the root's init, a commit or a memset
    DevBuf<std::uint32_t> ticket_, tasks_, groups_, taskCount_;
    ViewEngine view_;
};

// ---------------------------------------------------------------------------
// KDTree3D<Real> (src/kdtree3d/kdtree3d.cu)
// ---------------------------------------------------------------------------
template <class Real>
class BinaryImpl final : public TreeBase
{
public:
    BinaryImpl(const Real *x, const Real *y, const Real *z, std::size_t n, const at_config &c, unsigned maxDepth,
               unsigned iterations)
        : TreeBase(hipStream_t(c.stream)), view_(1)
    {
        x_ = x; y_ = y; z_ = z; n_ = n;
        maxDepth_ = maxDepth;
        bucket_ = c.max_particles_per_leaf;
        iterations_ = iterations;
        box_ = c.box;
        views_ = {c.views[0] != 0, c.views[1] != 0, c.views[2] != 0};
        prepared_ = views_;
        geom_ = makeBinaryGeom<Real>(box_, maxDepth_);
        P_.ensure(n ? n : 1, 0, stream_);
        perm_.ensure(n ? n : 1, 0, stream_);
        sort_.allocate(n ? n : 1, true, stream_);
        idx_.ensure(indexSize(n) + 1, 0, stream_);
        tileFirst_.ensure(n / kBoundTile + 2, 0, stream_);
        syncStream(stream_);
    }

    ~BinaryImpl() override { (void)syncStreamStatus(stream_); }

    void build() override
    {
        syncStream(stream_);
        built_ = false;
        hostLeaves_ = 1;
        prepareGraphBuffers(std::max<std::size_t>(2, leafCapacity_));
        computeParticleKeys();
        const BinParams p{bucket_, maxDepth_};
        if (bucket_ <= kDirectMaxBucket)
            buildDirect(p);
        else
            buildByPasses(p);
        built_ = true;
        prepareViews();
        buildTreeView(views_);
        if (views_.any()) (void)needsResize();
    }

    void update() override
    {
        if (!built_) throw Error(AT_ELOGIC, "adaptive_octree: call build() before update()");
        computeParticleKeys();
        const TreeDev t = dev();
        const BinParams p{bucket_, maxDepth_};
        updateReset(stream_, t);
        // the particles moved: count every leaf (small leaves: one read of P)
        if (bucket_ <= kTileCountBucket && n_ > 0) leafCountsByTiles(stream_, t);
        else leafCounts(stream_, t, leafCapacity_);
        for (unsigned i = 0; i < iterations_; ++i) binPass(stream_, t, p, long(leafCapacity_), leafCapacity_);
        buildTreeView(views_);
        AT_HIP(hipGetLastError());
    }

    void computeViews(const ViewFlags *v) override
    {
        if (!built_) throw Error(AT_ELOGIC, "adaptive_octree: call build() before computeViews()");
        buildTreeView(v ? *v : views_);
    }

    bool needsResize() override
    {
        if (!built_) return false;
        syncState();
        return sth_->needsResize != 0;
    }

    void resizeBuffers() override
    {
        if (!built_) throw Error(AT_ELOGIC, "adaptive_octree: call build() before resizeBuffers()");
        syncState();
        growBuffers();
        syncStream(stream_);
    }

    void buffer(int id, const void **d, std::size_t *s, std::size_t *c) override
    {
        const int ab = activeBuffer_;
        switch (id)
        {
        case AT_BUF_PARTICLE_KEYS: *d = P_.p; *s = *c = n_; return;
        case AT_BUF_PERM: *d = perm_.p; *s = *c = n_; return;
        case AT_BUF_CORNERSTONE: *d = K_[ab].p; *s = hostLeaves_ + 1; *c = leafCapacity_ + 1; return;
        case AT_BUF_LEAF_COUNTS: *d = N_[ab].p; *s = hostLeaves_; *c = leafCapacity_; return;
        case AT_BUF_NEAR_FIELD_COUNTS: throw Error(AT_ELOGIC, "Tree: Binary NFCount is not implemented yet");
        default: OctreeImpl<Real>::viewBuffer(view_, id, d, s, c);
        }
    }

    void scalars(at_scalars *o) override
    {
        o->num_particles = n_;
        o->num_active_particles = numActive_;
        o->num_nodes = numNodes_;
        o->num_leaves = hostLeaves_;
        o->leaf_capacity = leafCapacity_;
        o->tree_type = AT_TREE_BINARY;
        o->split_criterion = AT_LEAF_COUNT;
        o->sfc_kind = AT_MORTON;
        o->real_bytes = int(sizeof(Real));
        o->active_buffer = activeBuffer_;
        o->update_iterations = int(iterations_);
        o->max_depth = int(maxDepth_);
        o->max_particles_per_leaf = bucket_;
    }

    unsigned maxAchievedDepth() override // kdtree3d.cu:404-415
    {
        if (views_.any()) return view_.maxAchievedDepth(stream_);
        syncStream(stream_);
        std::vector<Key> k(hostLeaves_ + 1);
        AT_HIP(hipMemcpy(k.data(), K_[activeBuffer_].p, k.size() * sizeof(Key), hipMemcpyDeviceToHost));
        unsigned d = 0;
        for (std::size_t i = 0; i + 1 < k.size(); ++i) d = std::max(d, unsigned(clz64(k[i + 1] - k[i])));
        return d;
    }

private:
    TreeDev dev()
    {
        TreeDev t;
        for (int h = 0; h < 2; ++h)
        {
            t.K[h] = K_[h].p;
            t.N[h] = N_[h].p;
        }
        t.offs = offs_.p;
        t.scanStatus = status_.p;
        t.scanTicket = ticket_.p;
        t.statusWords = status_.cap;
        t.st = st_;
        t.P = P_.p;
        t.Pidx = idx_.p;
        t.tileFirst = tileFirst_.p;
        t.n = n_;
        t.capacity = long(leafCapacity_);
        return t;
    }

    void computeParticleKeys()
    {
        KeygenParams<Real> kp{};
        kp.x = x_; kp.y = y_; kp.z = z_;
        kp.geom = geom_;
        keygenAndSort<Real>(stream_, KeyKind::Binary, kp, n_, P_.p, perm_.p, sort_);
        buildSearchIndex(stream_, dev());
    }

    // The converged tree is unique (internal <=> count > bucket and depth <
    // maxDepth), so the reference's one-level-per-pass build loop
    // (kdtree3d.cu:240-252) is replaced by one direct construction; the
    // capacity it would reach grows by 25% steps from the starting capacity
    // until the final leaf count fits, the same rule as growBuffers.
    void buildDirect(const BinParams &p)
    {
        const std::size_t tiles = n_ / kPassTile + 2;
        direct_.delta.ensure(n_ ? n_ : 1, 0, stream_);
        direct_.cnt.ensure((n_ / kPassTile + 1) * kPassTile, 0, stream_); // whole tiles: 8-byte loads
        direct_.tileSum.ensure(tiles, 0, stream_);
        direct_.total.ensure(1, 0, stream_);
        direct_.ticket.ensure(1, 0, stream_);
        BinDirectScratch d;
        d.delta = direct_.delta.p;
        d.cnt = direct_.cnt.p;
        d.tileSum = direct_.tileSum.p;
        d.total = direct_.total.p;
        d.ticket = direct_.ticket.p;
        binDirectCount(stream_, dev(), p, d);
        // emitted at the current capacity; when the leaves do not fit, the
        // buffers grow and the emission runs again (the counts stay valid)
        direct_.bound.ensure(leafCapacity_ + 1, 0, stream_);
        d.bound = direct_.bound.p;
        binDirectWrite(stream_, dev(), p, d, long(leafCapacity_));
        syncState();
        const std::size_t L = std::size_t(sth_->numLeaves);
        if (L <= leafCapacity_) return;
        std::size_t cap = leafCapacity_;
        while (cap < L)
        {
            cap = growBy25(cap);
            constexpr std::size_t maxLeaves = std::size_t(std::numeric_limits<int>::max()) / 2 + 1;
            if (cap > maxLeaves) throw Error(AT_ELENGTH, "adaptive_octree: buffer capacity exceeds node index range");
        }
        hostLeaves_ = 0; // nothing was emitted: no live entries to keep
        prepareGraphBuffers(cap);
        direct_.bound.ensure(L + 1, 0, stream_);
        d.bound = direct_.bound.p;
        binDirectWrite(stream_, dev(), p, d, long(leafCapacity_));
        syncState();
        if (std::size_t(sth_->numLeaves) != L || L > leafCapacity_)
            throw Error(AT_EHIP, "binary construction produced an inconsistent leaf count");
    }

    // The reference loop (kdtree3d.cu:236-252), used for buckets whose
    // windows do not fit the LDS tile of the direct construction.
    void buildByPasses(const BinParams &p)
    {
        binInitRoot(stream_, dev());
        leafCounts(stream_, dev(), leafCapacity_);
        for (;;)
        {
            binPass(stream_, dev(), p, long(leafCapacity_), leafCapacity_);
            syncState();
            if (sth_->needsResize)
            {
                growBuffers();
                const TreeDev t = dev();
                updateReset(stream_, t);
                leafCounts(stream_, t, leafCapacity_);
            }
            else if (sth_->converged)
                break;
        }
    }

    void syncState() // readState (kdtree3d.cu:127-137)
    {
        readState();
        activeBuffer_ = sth_->activeBuffer;
        numActive_ = sth_->numActiveParticles;
        numNodes_ = std::size_t(sth_->numNodes);
        syncSizes(std::size_t(sth_->numLeaves));
    }
    void syncSizes(std::size_t leaves)
    {
        hostLeaves_ = leaves;
        view_.syncSizes(leaves, numActive_);
    }

    void prepareViews()
    {
        if (!prepared_.any()) return;
        view_.prepare(leafCapacity_, n_, prepared_.structure, prepared_.counts, prepared_.mapping, stream_);
    }

    // prepareGraphBuffers (kdtree3d.cu:96-116)
    void prepareGraphBuffers(std::size_t capacity)
    {
        constexpr std::size_t maxLeaves = std::size_t(std::numeric_limits<int>::max()) / 2 + 1;
        if (capacity > maxLeaves) throw Error(AT_ELENGTH, "adaptive_octree: buffer capacity exceeds node index range");
        leafCapacity_ = capacity;
        const int ab = activeBuffer_;
        K_[ab].ensure(capacity + 1, hostLeaves_ + 1, stream_);
        N_[ab].ensure(capacity, hostLeaves_, stream_);
        K_[ab ^ 1].ensure(capacity + 1, 0, stream_);
        N_[ab ^ 1].ensure(capacity, 0, stream_);
        offs_.ensure(capacity + 1, 0, stream_);
        status_.ensure(scanStatusWords(capacity + 1), 0, stream_);
        ticket_.ensure(1, 0, stream_);
        if (built_) prepareViews();
        syncSizes(hostLeaves_);
        syncStream(stream_);
    }

    void growBuffers()
    {
        prepareGraphBuffers(growBy25(leafCapacity_));
        AT_HIP(hipMemsetAsync(&st_->needsResize, 0, sizeof(int), stream_));
    }

    void buildTreeView(const ViewFlags &v)
    {
        prepared_.structure |= v.structure;
        prepared_.counts |= v.counts;
        prepared_.mapping |= v.mapping;
        Key *const K[2] = {K_[0].p, K_[1].p};
        std::uint32_t *const N[2] = {N_[0].p, N_[1].p};
        view_.build(stream_, K, N, P_.p, n_, v.structure, v.counts, v.mapping, st_);
    }

    struct DirectBufs
    {
        DevBuf<std::int8_t> delta;
        DevBuf<std::uint8_t> cnt;
        DevBuf<std::uint32_t> tileSum, bound, ticket;
        DevBuf<unsigned long long> total;
    };

    const Real *x_ = nullptr, *y_ = nullptr, *z_ = nullptr;
    std::size_t n_ = 0;
    unsigned maxDepth_ = 63, bucket_ = 64, iterations_ = 6;
    at_box box_{};
    BinaryGeom geom_{};
    ViewFlags views_, prepared_;
    bool built_ = false;
    int activeBuffer_ = 0;
    std::size_t leafCapacity_ = 0, hostLeaves_ = 1, numActive_ = 0, numNodes_ = 0;
    DevBuf<Key> P_;
    DevBuf<std::uint32_t> perm_;
    DevBuf<Key> K_[2];
    DevBuf<std::uint32_t> N_[2];
    DevBuf<std::int64_t> offs_;
    DevBuf<std::uint64_t> status_;
    DevBuf<std::uint32_t> ticket_;
    DevBuf<Key> idx_;
    DevBuf<std::uint32_t> tileFirst_; // per tile of P: first leaf boundary it owns (full leaf counts)
    DirectBufs direct_;
    ViewEngine view_;
};

template <class Real>
TreeBase *makeTree(int type, const Real *x, const Real *y, const Real *z, std::size_t n, const at_config &c)
{
    // Tree::backendConfig (tree.hpp:108-133) and the backend constructors
    if (c.max_depth < -1 || c.update_iterations < -1) throw Error(AT_EINVAL, "Tree: use -1 for backend defaults");
    if (type == AT_TREE_OCTREE)
    {
        if (n > 0 && (!x || !y || !z))
            throw Error(AT_EINVAL, "adaptive_octree: coordinate device pointers must not be null");
        if (c.max_particles_per_leaf == 0) throw Error(AT_EINVAL, "adaptive_octree: particle limit must be positive");
        const unsigned it = c.update_iterations >= 0 ? unsigned(c.update_iterations) : 2u;
        if (it == 0) throw Error(AT_EINVAL, "adaptive_octree: updateIterations must be positive");
        if (n > std::size_t(std::numeric_limits<std::uint32_t>::max()))
            throw Error(AT_EINVAL, "adaptive_octree: particle count exceeds 32-bit permutation index range");
        const unsigned md = c.max_depth >= 0 ? unsigned(c.max_depth) : kOctLevels;
        return new OctreeImpl<Real>(x, y, z, n, c, md, it);
    }
    if (type != AT_TREE_BINARY) throw Error(AT_EINVAL, "unknown tree type");
    if (c.split_criterion != AT_LEAF_COUNT) throw Error(AT_EINVAL, "Tree: Binary NFCount is not implemented yet");
    if (c.sfc_kind != AT_MORTON) throw Error(AT_EINVAL, "Tree: Binary uses longest-side keys, not Hilbert");
    if (n > 0 && (!x || !y || !z))
        throw Error(AT_EINVAL, "adaptive_octree: coordinate device pointers must not be null");
    if (c.max_particles_per_leaf == 0) throw Error(AT_EINVAL, "adaptive_octree: particle limit must be positive");
    const unsigned it = c.update_iterations >= 0 ? unsigned(c.update_iterations) : 6u;
    if (it == 0) throw Error(AT_EINVAL, "adaptive_octree: updateIterations must be positive");
    if (n > std::size_t(std::numeric_limits<std::uint32_t>::max()))
        throw Error(AT_EINVAL, "adaptive_octree: particle count exceeds 32-bit permutation index range");
    const unsigned md = c.max_depth >= 0 ? unsigned(c.max_depth) : kBinLevels;
    return new BinaryImpl<Real>(x, y, z, n, c, md, it);
}
} // namespace impl
} // namespace at
